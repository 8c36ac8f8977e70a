"""Python front end of libtmvp.so, the MI355X Saber/Sable TMVP library.

Everything here is plumbing around the C ABI declared in ``include/tmvp.h``:
the products themselves run in the HIP kernels of ``libtmvp.so``.  Device
memory, streams and events come straight from the HIP runtime
(``libamdhip64``) through ctypes; PyTorch is not needed.  If ``libtmvp.so`` has
not been built, or no gfx950 device is present, every entry point raises -
there is no CPU fallback.

Reference correspondence (paths relative to the reference root):
  matvec()     MatVecMul_gpu_shared   TMVP-Saber/SABER-GPU/src/poly.cu:269-376
  innerprod()  VecVecMul_gpu_shared   TMVP-Saber/SABER-GPU/src/poly.cu:227-248
               VecVecMul_Inner_gpu    TMVP_SABLE/src/pack.cu:262-286
  polymul()    one TMVP (convertnegacyclictest .. matvecout_cuda),
               TMVP-Saber/SABER-GPU/src/poly.cu:1116-1392
"""
import ctypes
import os
import subprocess

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libtmvp.so")
HIP_RUNTIME = os.environ.get("TMVP_HIP_RUNTIME", "/opt/rocm/lib/libamdhip64.so.7")

N = 256

# parameter sets (include/tmvp.h)
LIGHTSABER, SABER, FIRESABER, SABLE1, SABLE2, SABLE3 = range(6)
PARAM_SET_NAMES = {"lightsaber": 0, "saber": 1, "firesaber": 2, "sable1": 3, "sable2": 4, "sable3": 5}
# backends: reference mode 0 -> MFMA, 1 -> VALU, 2 -> schoolbook (VALU)
MFMA, VALU, SCHOOLBOOK = 0, 1, 2
ACCUMULATE = 0x1
# MFMA tile flags (include/tmvp.h)
TILE_16, TILE_32 = 0x2, 0x4
# Wrapper convenience for the operator calls below: the MFMA backend with the
# 32x32x32 int8 tile (passed to the library as MFMA + TILE_32).
MFMA32 = 16


IP_RAW, IP_ENC, IP_DEC = 0, 1, 2


def _backend_flags(backend, flags):
    return (MFMA, flags | TILE_32) if backend == MFMA32 else (backend, flags)

# status codes (include/tmvp.h)
OK, EARG, EHIP, ENOMEM, EUNSUPPORTED, ERANGE, ENODEV = 0, -1, -2, -3, -4, -5, -6
ERRORS = {0: "ok", -1: "invalid argument", -2: "HIP runtime error", -3: "out of memory",
          -4: "unsupported by this backend", -5: "secret coefficient outside the int8 MFMA range",
          -6: "no gfx950 device"}


class TmvpError(RuntimeError):
    def __init__(self, code, where=""):
        self.code = code
        super().__init__(f"{where}: {ERRORS.get(code, 'unknown error')} ({code})")


def build(force=True):
    """Compile libtmvp.so in place with hipcc (gfx950)."""
    cmd = ["make", "-C", PKG_DIR, "-j4"]
    if force:
        cmd.append("-B")
    subprocess.run(cmd, check=True)


class _Params(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in
                ("l", "eq", "ep", "et", "mu", "sable", "h1", "h2", "pk_bytes", "sk_bytes", "ct_bytes")]


_hip = None
_lib = None

# The exported C-ABI symbols, as declared in include/tmvp.h.
EXPORTS = ("tmvp_init", "tmvp_destroy", "tmvp_ctx_status", "tmvp_strerror", "tmvp_get_params",
           "tmvp_version", "tmvp_polymul", "tmvp_matvec", "tmvp_innerprod",
           "tmvp_kem_keypair", "tmvp_kem_encaps", "tmvp_kem_decaps",
           "tmvp_host_configure", "tmvp_host_set_randomness", "tmvp_host_status", "tmvp_host_release",
           "saber_enc", "saber_dec", "SABLE_enc", "SABLE_dec",
           "tmvp_batch_launch", "tmvp_batch_autotune", "tmvp_batch_plan", "tmvp_graph_begin", "tmvp_graph_end", "tmvp_graph_launch",
           "tmvp_graph_destroy", "tmvp_kem_reserve", "tmvp_kem_encaps_noalloc",
           "tmvp_srv_create", "tmvp_srv_add_keys", "tmvp_srv_submit", "tmvp_srv_poll", "tmvp_srv_drain",
           "tmvp_srv_get_stats", "tmvp_srv_destroy")

SRV_ENCAPS, SRV_DECAPS = 0, 1


class SrvConfig(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_int64), ("max_batch", ctypes.c_int64), ("slots", ctypes.c_int),
                ("backend", ctypes.c_int)]


class SrvRequest(ctypes.Structure):
    _fields_ = [("op", ctypes.c_int), ("param_set", ctypes.c_int), ("count", ctypes.c_int64),
                ("keys", ctypes.c_void_p), ("in_", ctypes.c_void_p), ("ct", ctypes.c_void_p),
                ("ss", ctypes.c_void_p), ("fail", ctypes.c_void_p), ("tag", ctypes.c_uint64)]


class SrvCompletion(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("tag", ctypes.c_uint64), ("status", ctypes.c_int),
                ("batch_requests", ctypes.c_int), ("batch_instances", ctypes.c_int64),
                ("submit_us", ctypes.c_double), ("start_us", ctypes.c_double), ("done_us", ctypes.c_double)]


class SrvStats(ctypes.Structure):
    _fields_ = [("requests_submitted", ctypes.c_int64), ("requests_done", ctypes.c_int64),
                ("instances_done", ctypes.c_int64), ("batches", ctypes.c_int64), ("slots", ctypes.c_int64),
                ("slot_bytes", ctypes.c_int64), ("max_concurrent", ctypes.c_int64), ("hbm_budget", ctypes.c_int64),
                ("store_bytes", ctypes.c_int64), ("arena_bytes", ctypes.c_int64),
                ("max_inflight_bytes", ctypes.c_int64)]

OP_POLYMUL, OP_MATVEC, OP_MATVEC_T, OP_INNERPROD = 0, 1, 2, 3
MAX_WORK = 16


class WorkDesc(ctypes.Structure):
    """tmvp_work_desc: one job of a mixed batch launch."""
    _fields_ = [("op", ctypes.c_int), ("param_set", ctypes.c_int), ("epilogue", ctypes.c_int),
                ("eq_bits", ctypes.c_int), ("flags", ctypes.c_int), ("batch", ctypes.c_int64),
                ("out", ctypes.c_void_p), ("a", ctypes.c_void_p), ("s", ctypes.c_void_p), ("aux", ctypes.c_void_p)]


def load_library(path=LIB_PATH):
    """Load libtmvp.so (and the HIP runtime it links).  Raises if missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built; run tmvp_amd.build() (hipcc, gfx950)")
    lib = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
    P, i, i64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64
    sigs = {
        "tmvp_init": (i, [ctypes.POINTER(P), i]),
        "tmvp_destroy": (i, [P]),
        "tmvp_ctx_status": (i, [P, P]),
        "tmvp_strerror": (ctypes.c_char_p, [i]),
        "tmvp_get_params": (i, [i, ctypes.POINTER(_Params)]),
        "tmvp_version": (i, []),
        "tmvp_polymul": (i, [P, i, P, P, P, i64, i, i, P]),
        "tmvp_matvec": (i, [P, i, i, i, P, P, P, i64, i, i, P]),
        "tmvp_innerprod": (i, [P, i, i, i, P, P, P, P, i64, i, P]),
        "tmvp_kem_keypair": (i, [P, i, i, P, P, P, i64, P]),
        "tmvp_kem_encaps": (i, [P, i, i, P, P, P, P, i64, P]),
        "tmvp_kem_decaps": (i, [P, i, i, P, P, P, P, i64, P]),
        "tmvp_host_configure": (i, [i, i64, i, i64, i]),
        "tmvp_host_set_randomness": (i, [P, i64]),
        "tmvp_host_status": (i, []),
        "tmvp_host_release": (i, []),
        "saber_enc": (None, [ctypes.c_uint8, P, P, P]),
        "saber_dec": (None, [ctypes.c_uint8, P, P, P, P]),
        "SABLE_enc": (None, [P, P, P, i]),
        "SABLE_dec": (None, [P, P, P, P, i]),
        "tmvp_batch_launch": (i, [P, i, P, i, P]),
        "tmvp_batch_autotune": (i, [P, P, i, P, ctypes.POINTER(ctypes.c_float)]),
        "tmvp_batch_plan": (i, [P, P, i, ctypes.POINTER(i), ctypes.POINTER(i)]),
        "tmvp_graph_begin": (i, [P, P]),
        "tmvp_graph_end": (i, [P, P, ctypes.POINTER(P)]),
        "tmvp_graph_launch": (i, [P, P]),
        "tmvp_graph_destroy": (i, [P]),
        "tmvp_kem_reserve": (i, [P, i, i64, P]),
        "tmvp_kem_encaps_noalloc": (i, [P, i, i, P, P, P, P, i64, P]),
        "tmvp_srv_create": (i, [P, ctypes.POINTER(SrvConfig), ctypes.POINTER(P)]),
        "tmvp_srv_add_keys": (i, [P, i, P, P, i64, ctypes.POINTER(i64)]),
        "tmvp_srv_submit": (i, [P, ctypes.POINTER(SrvRequest), ctypes.POINTER(ctypes.c_uint64)]),
        "tmvp_srv_poll": (i, [P, P, i, i]),
        "tmvp_srv_drain": (i, [P]),
        "tmvp_srv_get_stats": (i, [P, ctypes.POINTER(SrvStats)]),
        "tmvp_srv_destroy": (i, [P]),
    }
    # Every export needs its signature: without one ctypes passes 64-bit
    # arguments as 32-bit ints (pointers truncated, int64 sizes with garbage
    # upper halves when they go on the stack).
    missing = set(EXPORTS) - set(sigs)
    if missing:
        raise RuntimeError(f"no ctypes signature for {sorted(missing)}")
    for name, (res, args) in sigs.items():
        f = getattr(lib, name)
        f.restype, f.argtypes = res, args
    _lib = lib
    return lib


def hip():
    """ctypes handle of the HIP runtime (the same one libtmvp.so links)."""
    global _hip
    if _hip is None:
        load_library()
        h = ctypes.CDLL(HIP_RUNTIME, mode=ctypes.RTLD_GLOBAL)
        P, i, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t
        sigs = {
            "hipMalloc": (i, [ctypes.POINTER(P), sz]),
            "hipFree": (i, [P]),
            "hipMemcpy": (i, [P, P, sz, i]),
            "hipMemcpyAsync": (i, [P, P, sz, i, P]),
            "hipMemset": (i, [P, i, sz]),
            "hipMemsetAsync": (i, [P, i, sz, P]),
            "hipDeviceSynchronize": (i, []),
            "hipStreamCreate": (i, [ctypes.POINTER(P)]),
            "hipHostMalloc": (i, [ctypes.POINTER(P), sz, ctypes.c_uint]),
            "hipHostFree": (i, [P]),
            "hipStreamDestroy": (i, [P]),
            "hipStreamSynchronize": (i, [P]),
            "hipEventCreate": (i, [ctypes.POINTER(P)]),
            "hipEventDestroy": (i, [P]),
            "hipEventRecord": (i, [P, P]),
            "hipEventSynchronize": (i, [P]),
            "hipEventElapsedTime": (i, [ctypes.POINTER(ctypes.c_float), P, P]),
            "hipGetDeviceCount": (i, [ctypes.POINTER(i)]),
            "hipSetDevice": (i, [i]),
            "hipGetErrorString": (ctypes.c_char_p, [i]),
        }
        for name, (res, args) in sigs.items():
            f = getattr(h, name)
            f.restype, f.argtypes = res, args
        _hip = h
    return _hip


def _hipcheck(rc, what):
    if rc != 0:
        msg = hip().hipGetErrorString(rc)
        raise RuntimeError(f"{what} failed: {msg.decode() if msg else rc}")


H2D, D2H, D2D = 1, 2, 3


class DeviceBuffer:
    """A hipMalloc'd buffer owned by Python."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = ctypes.c_void_p()
        _hipcheck(hip().hipMalloc(ctypes.byref(self.ptr), max(self.nbytes, 1)), "hipMalloc")

    @classmethod
    def from_numpy(cls, arr):
        arr = np.ascontiguousarray(arr)
        buf = cls(arr.nbytes)
        buf.upload(arr)
        return buf

    def upload(self, arr, stream=None):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        if stream is None:
            _hipcheck(hip().hipMemcpy(self.ptr, arr.ctypes.data_as(ctypes.c_void_p), arr.nbytes, H2D), "hipMemcpy")
        else:
            _hipcheck(hip().hipMemcpyAsync(self.ptr, arr.ctypes.data_as(ctypes.c_void_p), arr.nbytes, H2D,
                                           _stream(stream)), "hipMemcpyAsync")
        return self

    def download(self, dtype=np.uint8, shape=None):
        out = np.empty(self.nbytes // np.dtype(dtype).itemsize, dtype)
        _hipcheck(hip().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, out.nbytes, D2H), "hipMemcpy")
        return out if shape is None else out.reshape(shape)

    def zero(self, stream=None):
        if stream is None:
            _hipcheck(hip().hipMemset(self.ptr, 0, self.nbytes), "hipMemset")
        else:
            _hipcheck(hip().hipMemsetAsync(self.ptr, 0, self.nbytes, _stream(stream)), "hipMemsetAsync")
        return self

    def free(self):
        if self.ptr is not None and self.ptr.value:
            hip().hipFree(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        self.handle = ctypes.c_void_p()
        _hipcheck(hip().hipStreamCreate(ctypes.byref(self.handle)), "hipStreamCreate")

    def synchronize(self):
        _hipcheck(hip().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def destroy(self):
        if self.handle is not None:
            hip().hipStreamDestroy(self.handle)
            self.handle = None


class Event:
    def __init__(self):
        self.handle = ctypes.c_void_p()
        _hipcheck(hip().hipEventCreate(ctypes.byref(self.handle)), "hipEventCreate")

    def record(self, stream=None):
        _hipcheck(hip().hipEventRecord(self.handle, _stream(stream)), "hipEventRecord")

    def synchronize(self):
        _hipcheck(hip().hipEventSynchronize(self.handle), "hipEventSynchronize")

    def elapsed_ms(self, end):
        ms = ctypes.c_float()
        _hipcheck(hip().hipEventElapsedTime(ctypes.byref(ms), self.handle, end.handle), "hipEventElapsedTime")
        return ms.value

    def destroy(self):
        if self.handle is not None:
            hip().hipEventDestroy(self.handle)
            self.handle = None


def device_synchronize():
    _hipcheck(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


def get_params(ps):
    ps = PARAM_SET_NAMES.get(ps, ps)
    out = _Params()
    rc = load_library().tmvp_get_params(int(ps), ctypes.byref(out))
    if rc:
        raise TmvpError(rc, "tmvp_get_params")
    return {n: getattr(out, n) for n, _ in _Params._fields_}


def _stream(st):
    if isinstance(st, Stream):
        return st.handle
    return st


def _ptr(x):
    if isinstance(x, DeviceBuffer):
        return x.ptr
    if x is None:
        return None
    return ctypes.c_void_p(int(x))


class Context:
    """A tmvp_ctx on one device.  Methods mirror the C ABI one to one."""

    def __init__(self, device=0):
        lib = load_library()
        hip()
        self.handle = ctypes.c_void_p()
        rc = lib.tmvp_init(ctypes.byref(self.handle), int(device))
        if rc:
            raise TmvpError(rc, "tmvp_init")

    def close(self):
        if self.handle is not None and self.handle.value:
            load_library().tmvp_destroy(self.handle)
        self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def status(self, stream=None):
        return load_library().tmvp_ctx_status(self.handle, _stream(stream))

    def check(self, stream=None):
        rc = self.status(stream)
        if rc:
            raise TmvpError(rc, "tmvp_ctx_status")

    def polymul(self, c, a, s, batch, eq_bits=13, backend=MFMA, flags=0, stream=None):
        backend, flags = _backend_flags(backend, flags)
        rc = load_library().tmvp_polymul(self.handle, backend, _ptr(c), _ptr(a), _ptr(s), int(batch),
                                         int(eq_bits), int(flags), _stream(stream))
        if rc:
            raise TmvpError(rc, "tmvp_polymul")

    def matvec(self, out, A, s, batch, param_set=SABER, backend=MFMA, transpose=False, round_to_p=False,
               flags=0, stream=None):
        backend, flags = _backend_flags(backend, flags)
        ps = PARAM_SET_NAMES.get(param_set, param_set)
        rc = load_library().tmvp_matvec(self.handle, int(ps), backend, int(bool(transpose)), _ptr(out), _ptr(A),
                                        _ptr(s), int(batch), int(bool(round_to_p)), int(flags), _stream(stream))
        if rc:
            raise TmvpError(rc, "tmvp_matvec")

    def innerprod(self, out, b, s, batch, param_set=SABER, backend=MFMA, epilogue=IP_RAW, aux=None, flags=0,
                  stream=None):
        backend, flags = _backend_flags(backend, flags)
        ps = PARAM_SET_NAMES.get(param_set, param_set)
        rc = load_library().tmvp_innerprod(self.handle, int(ps), backend, int(epilogue), _ptr(out), _ptr(b),
                                           _ptr(s), _ptr(aux), int(batch), int(flags), _stream(stream))
        if rc:
            raise TmvpError(rc, "tmvp_innerprod")


    def kem_keypair(self, pk, sk, seeds, batch, param_set=SABER, backend=MFMA, stream=None):
        ps = PARAM_SET_NAMES.get(param_set, param_set)
        rc = load_library().tmvp_kem_keypair(self.handle, int(ps), backend, _ptr(pk), _ptr(sk), _ptr(seeds),
                                             int(batch), _stream(stream))
        if rc:
            raise TmvpError(rc, "tmvp_kem_keypair")

    def kem_encaps(self, ct, ss, pk, rnd, batch, param_set=SABER, backend=MFMA, stream=None):
        ps = PARAM_SET_NAMES.get(param_set, param_set)
        rc = load_library().tmvp_kem_encaps(self.handle, int(ps), backend, _ptr(ct), _ptr(ss), _ptr(pk), _ptr(rnd),
                                            int(batch), _stream(stream))
        if rc:
            raise TmvpError(rc, "tmvp_kem_encaps")

    def kem_decaps(self, ss, ct, sk, batch, param_set=SABER, backend=MFMA, fail=None, stream=None):
        ps = PARAM_SET_NAMES.get(param_set, param_set)
        rc = load_library().tmvp_kem_decaps(self.handle, int(ps), backend, _ptr(ss), _ptr(ct), _ptr(sk),
                                            _ptr(fail), int(batch), _stream(stream))
        if rc:
            raise TmvpError(rc, "tmvp_kem_decaps")


    def batch_launch(self, descs, backend=MFMA, stream=None):
        arr = (WorkDesc * max(1, len(descs)))(*descs)
        rc = load_library().tmvp_batch_launch(self.handle, backend, ctypes.cast(arr, ctypes.c_void_p), len(descs),
                                              _stream(stream))
        if rc:
            raise TmvpError(rc, "tmvp_batch_launch")

    def batch_autotune(self, descs, stream=None):
        arr = (WorkDesc * max(1, len(descs)))(*descs)
        ms = ctypes.c_float()
        rc = load_library().tmvp_batch_autotune(self.handle, ctypes.cast(arr, ctypes.c_void_p), len(descs),
                                                _stream(stream), ctypes.byref(ms))
        if rc:
            raise TmvpError(rc, "tmvp_batch_autotune")
        return ms.value

    def batch_plan(self, descs):
        """(tile shape, workgroups per CU, measured) of the plan this list uses."""
        arr = (WorkDesc * len(descs))(*descs)
        shape, wgs = ctypes.c_int(), ctypes.c_int()
        rc = load_library().tmvp_batch_plan(self.handle, ctypes.cast(arr, ctypes.c_void_p), len(descs),
                                            ctypes.byref(shape), ctypes.byref(wgs))
        if rc < 0:
            raise TmvpError(rc, "tmvp_batch_plan")
        return shape.value, wgs.value, rc == 0

    def kem_reserve(self, param_set, batch, stream=None):
        ps = PARAM_SET_NAMES.get(param_set, param_set)
        rc = load_library().tmvp_kem_reserve(self.handle, int(ps), int(batch), _stream(stream))
        if rc:
            raise TmvpError(rc, "tmvp_kem_reserve")

    def capture(self, stream, fn):
        """Record fn() (library calls on `stream`) into a Graph."""
        lib = load_library()
        rc = lib.tmvp_graph_begin(self.handle, _stream(stream))
        if rc:
            raise TmvpError(rc, "tmvp_graph_begin")
        try:
            fn()
        finally:
            g = ctypes.c_void_p()
            rc = lib.tmvp_graph_end(self.handle, _stream(stream), ctypes.byref(g))
        if rc:
            raise TmvpError(rc, "tmvp_graph_end")
        return Graph(g)


class Graph:
    """An instantiated hipGraph of library calls (tmvp_graph)."""

    def __init__(self, handle):
        self.handle = handle

    def launch(self, stream=None):
        rc = load_library().tmvp_graph_launch(self.handle, _stream(stream))
        if rc:
            raise TmvpError(rc, "tmvp_graph_launch")

    def destroy(self):
        if self.handle is not None and self.handle.value:
            load_library().tmvp_graph_destroy(self.handle)
        self.handle = None


def work(op, out, a, s, batch, param_set=SABER, epilogue=0, aux=None, eq_bits=0, flags=0):
    """A WorkDesc over DeviceBuffers (or raw device addresses)."""
    return WorkDesc(op, int(PARAM_SET_NAMES.get(param_set, param_set)), epilogue, eq_bits, flags, int(batch),
                    _ptr(out), _ptr(a), _ptr(s), _ptr(aux))


# ---- numpy conveniences (allocate, copy, run, copy back, synchronise) ----

def run_polymul(ctx, a, s, eq_bits=13, backend=MFMA):
    a = np.ascontiguousarray(a, np.uint16).reshape(-1, N)
    s = np.ascontiguousarray(s, np.uint16).reshape(-1, N)
    batch = a.shape[0]
    da, ds_, dc = DeviceBuffer.from_numpy(a), DeviceBuffer.from_numpy(s), DeviceBuffer(a.nbytes)
    ctx.polymul(dc, da, ds_, batch, eq_bits, backend)
    device_synchronize()
    ctx.check()
    return dc.download(np.uint16, (batch, N))


def run_matvec(ctx, A, s, param_set=SABER, backend=MFMA, transpose=False, round_to_p=False, out_init=None):
    l = get_params(param_set)["l"]
    A = np.ascontiguousarray(A, np.uint16).reshape(-1, l * l * N)
    s = np.ascontiguousarray(s, np.uint16).reshape(-1, l * N)
    batch = A.shape[0]
    dA, ds_ = DeviceBuffer.from_numpy(A), DeviceBuffer.from_numpy(s)
    dout = DeviceBuffer(batch * l * N * 2)
    flags = 0
    if out_init is not None:
        dout.upload(np.ascontiguousarray(out_init, np.uint16))
        flags = ACCUMULATE
    ctx.matvec(dout, dA, ds_, batch, param_set, backend, transpose, round_to_p, flags)
    device_synchronize()
    ctx.check()
    return dout.download(np.uint16, (batch, l, N))


def run_innerprod(ctx, b, s, param_set=SABER, backend=MFMA, epilogue=IP_RAW, aux=None):
    l = get_params(param_set)["l"]
    b = np.ascontiguousarray(b, np.uint16).reshape(-1, l * N)
    s = np.ascontiguousarray(s, np.uint16).reshape(-1, l * N)
    batch = b.shape[0]
    db, ds_ = DeviceBuffer.from_numpy(b), DeviceBuffer.from_numpy(s)
    daux = DeviceBuffer.from_numpy(np.ascontiguousarray(aux, np.uint16)) if aux is not None else None
    dout = DeviceBuffer(batch * N * 2)
    ctx.innerprod(dout, db, ds_, batch, param_set, backend, epilogue, daux)
    device_synchronize()
    ctx.check()
    return dout.download(np.uint16, (batch, N))


def run_kem_keypair(ctx, seeds, param_set=SABER, backend=MFMA):
    p = get_params(param_set)
    seeds = np.ascontiguousarray(seeds, np.uint8).reshape(-1, 96)
    batch = seeds.shape[0]
    dseeds = DeviceBuffer.from_numpy(seeds)
    dpk, dsk = DeviceBuffer(batch * p["pk_bytes"]), DeviceBuffer(batch * p["sk_bytes"])
    ctx.kem_keypair(dpk, dsk, dseeds, batch, param_set, backend)
    device_synchronize()
    ctx.check()
    return dpk.download(np.uint8, (batch, -1)), dsk.download(np.uint8, (batch, -1))


def run_kem_encaps(ctx, pk, rnd, param_set=SABER, backend=MFMA):
    p = get_params(param_set)
    pk = np.ascontiguousarray(pk, np.uint8).reshape(-1, p["pk_bytes"])
    rnd = np.ascontiguousarray(rnd, np.uint8).reshape(-1, 32)
    batch = pk.shape[0]
    dpk, drnd = DeviceBuffer.from_numpy(pk), DeviceBuffer.from_numpy(rnd)
    dct, dss = DeviceBuffer(batch * p["ct_bytes"]), DeviceBuffer(batch * 32)
    ctx.kem_encaps(dct, dss, dpk, drnd, batch, param_set, backend)
    device_synchronize()
    ctx.check()
    return dct.download(np.uint8, (batch, -1)), dss.download(np.uint8, (batch, -1))


def run_kem_decaps(ctx, ct, sk, param_set=SABER, backend=MFMA):
    p = get_params(param_set)
    ct = np.ascontiguousarray(ct, np.uint8).reshape(-1, p["ct_bytes"])
    sk = np.ascontiguousarray(sk, np.uint8).reshape(-1, p["sk_bytes"])
    batch = ct.shape[0]
    dct, dsk = DeviceBuffer.from_numpy(ct), DeviceBuffer.from_numpy(sk)
    dss, dfail = DeviceBuffer(batch * 32), DeviceBuffer(batch)
    ctx.kem_decaps(dss, dct, dsk, batch, param_set, backend, fail=dfail)
    device_synchronize()
    ctx.check()
    return dss.download(np.uint8, (batch, -1)), dfail.download(np.uint8)


class _PinnedBlock:
    def __init__(self, nbytes):
        self.ptr = ctypes.c_void_p()
        _hipcheck(hip().hipHostMalloc(ctypes.byref(self.ptr), max(1, nbytes), 0), "hipHostMalloc")
        self.nbytes = nbytes

    def __del__(self):
        if self.ptr is not None and self.ptr.value:
            hip().hipHostFree(self.ptr)
            self.ptr = None


def pinned_array(shape, dtype=np.uint8):
    """A numpy array in page-locked host memory (hipHostMalloc); the memory is
    freed when the array and every view of it are gone."""
    dtype = np.dtype(dtype)
    nbytes = int(np.prod(shape)) * dtype.itemsize
    blk = _PinnedBlock(nbytes)
    raw = (ctypes.c_uint8 * max(1, nbytes)).from_address(blk.ptr.value)
    raw._owner = blk
    arr = np.frombuffer(raw, np.uint8, count=nbytes).view(dtype).reshape(shape)
    return arr


class Server:
    """Serving loop over device-resident keys (tmvp_server).  Host arrays of a
    submitted request are kept alive here until its completion is polled."""

    def __init__(self, ctx, hbm_budget=0, max_batch=0, slots=0, backend=MFMA):
        self.lib = load_library()
        self.handle = ctypes.c_void_p()
        cfg = SrvConfig(hbm_budget, max_batch, slots, backend)
        rc = self.lib.tmvp_srv_create(ctx.handle, ctypes.byref(cfg), ctypes.byref(self.handle))
        if rc:
            raise TmvpError(rc, "tmvp_srv_create")
        self.live = {}

    def add_keys(self, param_set, pk, sk=None):
        ps = PARAM_SET_NAMES.get(param_set, param_set)
        pk = np.ascontiguousarray(pk, np.uint8)
        n = pk.shape[0]
        first = ctypes.c_int64()
        skp = None
        if sk is not None:
            sk = np.ascontiguousarray(sk, np.uint8)
            skp = sk.ctypes.data_as(ctypes.c_void_p)
        rc = self.lib.tmvp_srv_add_keys(self.handle, int(ps), pk.ctypes.data_as(ctypes.c_void_p), skp, n,
                                        ctypes.byref(first))
        if rc:
            raise TmvpError(rc, "tmvp_srv_add_keys")
        return first.value

    def submit(self, op, param_set, keys, data, tag=0, out=None):
        """Queue a request; returns (id, outputs) where outputs are the numpy
        arrays the results land in: (ct, ss) for encaps, (ss, fail) for decaps
        (`out` may supply them, e.g. views of pinned_array memory)."""
        ps = PARAM_SET_NAMES.get(param_set, param_set)
        p = get_params(ps)
        keys = np.ascontiguousarray(keys, np.int32)
        data = np.ascontiguousarray(data, np.uint8)
        n = keys.shape[0]
        if op == SRV_ENCAPS:
            ct, ss = out if out is not None else (np.zeros((n, p["ct_bytes"]), np.uint8), np.zeros((n, 32), np.uint8))
            ctp, failp = ct.ctypes.data_as(ctypes.c_void_p), None
            outs = (ct, ss)
        else:
            ss, fail = out if out is not None else (np.zeros((n, 32), np.uint8), np.zeros(n, np.uint8))
            ctp, failp = None, fail.ctypes.data_as(ctypes.c_void_p)
            outs = (ss, fail)
        req = SrvRequest(op, int(ps), n, keys.ctypes.data_as(ctypes.c_void_p), data.ctypes.data_as(ctypes.c_void_p),
                         ctp, ss.ctypes.data_as(ctypes.c_void_p), failp, tag)
        rid = ctypes.c_uint64()
        rc = self.lib.tmvp_srv_submit(self.handle, ctypes.byref(req), ctypes.byref(rid))
        if rc:
            raise TmvpError(rc, "tmvp_srv_submit")
        self.live[rid.value] = (keys, data, outs)
        return rid.value, outs

    def poll(self, max_n=256, timeout_us=0):
        buf = (SrvCompletion * max_n)()
        k = self.lib.tmvp_srv_poll(self.handle, ctypes.cast(buf, ctypes.c_void_p), max_n, int(timeout_us))
        if k < 0:
            raise TmvpError(k, "tmvp_srv_poll")
        out = []
        for c in buf[:k]:
            self.live.pop(c.id, None)
            out.append({"id": c.id, "tag": c.tag, "status": c.status, "batch_requests": c.batch_requests,
                        "batch_instances": c.batch_instances, "submit_us": c.submit_us, "start_us": c.start_us,
                        "done_us": c.done_us})
        return out

    def drain(self):
        rc = self.lib.tmvp_srv_drain(self.handle)
        if rc:
            raise TmvpError(rc, "tmvp_srv_drain")

    def stats(self):
        st = SrvStats()
        rc = self.lib.tmvp_srv_get_stats(self.handle, ctypes.byref(st))
        if rc:
            raise TmvpError(rc, "tmvp_srv_get_stats")
        return {f: getattr(st, f) for f, _ in SrvStats._fields_}

    def close(self):
        if self.handle is not None and self.handle.value:
            self.lib.tmvp_srv_destroy(self.handle)
        self.handle = None
