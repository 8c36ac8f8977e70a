"""ORACLE - test infrastructure only.

ctypes/numpy front end of ``liboracle.so``, the CPU restatement of the
reference's Saber/Sable arithmetic (see ``tmvp_oracle.c``).  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import this module, and only as the checker or the timed CPU baseline; the
GPU library never calls it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

N = 256
PARAM_SETS = {
    "lightsaber": 0, "saber": 1, "firesaber": 2,
    "sable1": 3, "sable2": 4, "sable3": 5,
}


class Params(ctypes.Structure):
    _fields_ = [(name, ctypes.c_int) for name in (
        "l", "eq", "ep", "et", "mu", "sable", "h1", "h2",
        "polyvec_compressed_bytes", "scale_bytes", "indcpa_pk_bytes",
        "indcpa_sk_bytes", "pk_bytes", "sk_bytes", "ct_bytes",
        "matrix_stream_bytes", "secret_stream_bytes")]


def build(force=False):
    """Compile liboracle.so with gcc (make)."""
    args = ["make", "-C", HERE]
    if force:
        args.append("-B")
    subprocess.run(args, check=True, stdout=subprocess.DEVNULL)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(L):
    P = ctypes.c_void_p
    i, i64, sz = ctypes.c_int, ctypes.c_int64, ctypes.c_size_t
    sig = {
        "or_get_params": (i, [i, ctypes.POINTER(Params)]),
        "or_shake128": (None, [P, sz, P, sz]),
        "or_sha3_256": (None, [P, P, sz]),
        "or_sha3_512": (None, [P, P, sz]),
        "or_polymul_schoolbook": (None, [P, P, P]),
        "or_polymul_toom4": (None, [P, P, P]),
        "or_matvec": (None, [i, P, P, P, i, i]),
        "or_innerprod": (None, [i, P, P, P, i]),
        "or_polymul_batch": (None, [P, P, P, i64, i, i]),
        "or_matvec_batch": (None, [i, P, P, P, i64, i, i, i]),
        "or_innerprod_batch": (None, [i, P, P, P, i64, i, i]),
        "or_bench_matvec1": (ctypes.c_double, [i, P, P, i, i]),
        "or_round_q2p": (ctypes.c_uint16, [i, ctypes.c_uint16]),
        "or_round_enc_v": (ctypes.c_uint16, [i, ctypes.c_uint16, ctypes.c_uint16]),
        "or_round_dec_m": (ctypes.c_uint16, [i, ctypes.c_uint16, ctypes.c_uint16]),
        "or_pack_bits": (None, [P, i, i, P]),
        "or_unpack_bits": (None, [P, i, i, P]),
        "or_cbd": (None, [i, P, P]),
        "or_gen_matrix": (None, [i, P, P]),
        "or_gen_secret": (None, [i, P, P]),
        "or_indcpa_keypair": (None, [i, P, P, P, P]),
        "or_indcpa_enc": (None, [i, P, P, P, P]),
        "or_indcpa_dec": (None, [i, P, P, P]),
        "or_kem_keypair": (None, [i, P, P, P, P, P]),
        "or_kem_enc": (None, [i, P, P, P, P]),
        "or_kem_dec": (i, [i, P, P, P]),
        "or_kem_keypair_batch": (None, [i, P, P, P, i64, i]),
        "or_kem_enc_batch": (None, [i, P, P, P, P, i64, i]),
        "or_kem_dec_batch": (None, [i, P, P, P, i64, i]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _ps(ps):
    return PARAM_SETS[ps] if isinstance(ps, str) else int(ps)


def params(ps):
    out = Params()
    if lib().or_get_params(_ps(ps), ctypes.byref(out)) != 0:
        raise ValueError(f"unknown parameter set {ps!r}")
    return {name: getattr(out, name) for name, _ in Params._fields_}


def _u16(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.uint16))


def _u8(x):
    if isinstance(x, (bytes, bytearray)):
        x = np.frombuffer(bytes(x), dtype=np.uint8)
    return np.ascontiguousarray(np.asarray(x, dtype=np.uint8))


# ---- hashing ----
def shake128(data, outlen):
    d = _u8(data)
    out = np.zeros(outlen, np.uint8)
    lib().or_shake128(_p(out), outlen, _p(d), d.size)
    return out.tobytes()


def sha3_256(data):
    d = _u8(data)
    out = np.zeros(32, np.uint8)
    lib().or_sha3_256(_p(out), _p(d), d.size)
    return out.tobytes()


def sha3_512(data):
    d = _u8(data)
    out = np.zeros(64, np.uint8)
    lib().or_sha3_512(_p(out), _p(d), d.size)
    return out.tobytes()


# ---- arithmetic (outputs are raw sums mod 2^16) ----
def polymul(a, s, method=0):
    a, s = _u16(a), _u16(s)
    c = np.zeros(N, np.uint16)
    (lib().or_polymul_toom4 if method == 1 else lib().or_polymul_schoolbook)(_p(a), _p(s), _p(c))
    return c


def polymul_batch(a, s, method=0, nthreads=1):
    a, s = _u16(a), _u16(s)
    batch = a.size // N
    c = np.zeros(batch * N, np.uint16)
    lib().or_polymul_batch(_p(a), _p(s), _p(c), batch, method, nthreads)
    return c.reshape(batch, N)


def matvec_batch(l, A, s, transpose=False, method=0, nthreads=1, out=None):
    A, s = _u16(A), _u16(s)
    batch = s.size // (l * N)
    if out is None:
        out = np.zeros(batch * l * N, np.uint16)
    else:
        out = _u16(out).copy().reshape(-1)
    lib().or_matvec_batch(l, _p(A), _p(s), _p(out), batch, int(bool(transpose)), method, nthreads)
    return out.reshape(batch, l, N)


def innerprod_batch(l, b, s, method=0, nthreads=1, out=None):
    b, s = _u16(b), _u16(s)
    batch = s.size // (l * N)
    if out is None:
        out = np.zeros(batch * N, np.uint16)
    else:
        out = _u16(out).copy().reshape(-1)
    lib().or_innerprod_batch(l, _p(b), _p(s), _p(out), batch, method, nthreads)
    return out.reshape(batch, N)


def bench_matvec1(l, A, s, method=0, reps=1000):
    """Mean ns of one single-instance MatrixVectorMul timed inside the C
    oracle (one thread); used only as the CPU baseline's batch-1 timing."""
    A, s = _u16(A).reshape(-1), _u16(s).reshape(-1)
    return lib().or_bench_matvec1(l, _p(A), _p(s), method, reps)


def round_q2p(ps, x):
    f = lib().or_round_q2p
    ps = _ps(ps)
    return np.array([f(ps, int(v)) for v in np.asarray(x).reshape(-1)], np.uint16).reshape(np.shape(x))


def round_enc_v(ps, v, m):
    f = lib().or_round_enc_v
    ps = _ps(ps)
    v, m = np.asarray(v).reshape(-1), np.asarray(m).reshape(-1)
    return np.array([f(ps, int(a), int(b)) for a, b in zip(v, m)], np.uint16)


def round_dec_m(ps, v, c):
    f = lib().or_round_dec_m
    ps = _ps(ps)
    v, c = np.asarray(v).reshape(-1), np.asarray(c).reshape(-1)
    return np.array([f(ps, int(a), int(b)) for a, b in zip(v, c)], np.uint16)


# ---- codecs / sampling ----
def pack_bits(data, w):
    d = _u16(data)
    out = np.zeros((d.size * w + 7) // 8, np.uint8)
    lib().or_pack_bits(_p(d), d.size, w, _p(out))
    return out.tobytes()


def unpack_bits(data, n, w):
    d = _u8(data)
    out = np.zeros(n, np.uint16)
    lib().or_unpack_bits(_p(d), n, w, _p(out))
    return out


def cbd(ps, buf):
    b = _u8(buf)
    out = np.zeros(N, np.uint16)
    lib().or_cbd(_ps(ps), _p(b), _p(out))
    return out


def gen_matrix(ps, seed):
    p = params(ps)
    sd = _u8(seed)
    out = np.zeros(p["l"] * p["l"] * N, np.uint16)
    lib().or_gen_matrix(_ps(ps), _p(sd), _p(out))
    return out


def gen_secret(ps, seed):
    p = params(ps)
    sd = _u8(seed)
    out = np.zeros(p["l"] * N, np.uint16)
    lib().or_gen_secret(_ps(ps), _p(sd), _p(out))
    return out


# ---- IND-CPA / KEM ----
def indcpa_keypair(ps, seed_a, seed_s):
    p = params(ps)
    pk = np.zeros(p["indcpa_pk_bytes"], np.uint8)
    sk = np.zeros(p["indcpa_sk_bytes"], np.uint8)
    lib().or_indcpa_keypair(_ps(ps), _p(_u8(seed_a)), _p(_u8(seed_s)), _p(pk), _p(sk))
    return pk.tobytes(), sk.tobytes()


def indcpa_enc(ps, m, coins, pk):
    p = params(ps)
    ct = np.zeros(p["ct_bytes"], np.uint8)
    lib().or_indcpa_enc(_ps(ps), _p(_u8(m)), _p(_u8(coins)), _p(_u8(pk)), _p(ct))
    return ct.tobytes()


def indcpa_dec(ps, sk, ct):
    m = np.zeros(32, np.uint8)
    lib().or_indcpa_dec(_ps(ps), _p(_u8(sk)), _p(_u8(ct)), _p(m))
    return m.tobytes()


def kem_keypair(ps, seed_a, seed_s, z):
    p = params(ps)
    pk = np.zeros(p["pk_bytes"], np.uint8)
    sk = np.zeros(p["sk_bytes"], np.uint8)
    lib().or_kem_keypair(_ps(ps), _p(_u8(seed_a)), _p(_u8(seed_s)), _p(_u8(z)), _p(pk), _p(sk))
    return pk.tobytes(), sk.tobytes()


def kem_enc(ps, rnd, pk):
    p = params(ps)
    ct = np.zeros(p["ct_bytes"], np.uint8)
    ss = np.zeros(32, np.uint8)
    lib().or_kem_enc(_ps(ps), _p(_u8(rnd)), _p(_u8(pk)), _p(ct), _p(ss))
    return ct.tobytes(), ss.tobytes()


def kem_dec(ps, ct, sk):
    ss = np.zeros(32, np.uint8)
    fail = lib().or_kem_dec(_ps(ps), _p(_u8(ct)), _p(_u8(sk)), _p(ss))
    return ss.tobytes(), bool(fail)


def kem_keypair_batch(ps, seeds, nthreads=1):
    p = params(ps)
    seeds = _u8(seeds)
    batch = seeds.size // 96
    pk = np.zeros(batch * p["pk_bytes"], np.uint8)
    sk = np.zeros(batch * p["sk_bytes"], np.uint8)
    lib().or_kem_keypair_batch(_ps(ps), _p(seeds), _p(pk), _p(sk), batch, nthreads)
    return pk.reshape(batch, -1), sk.reshape(batch, -1)


def kem_enc_batch(ps, rnd, pk, nthreads=1):
    p = params(ps)
    rnd, pk = _u8(rnd), _u8(pk)
    batch = rnd.size // 32
    ct = np.zeros(batch * p["ct_bytes"], np.uint8)
    ss = np.zeros(batch * 32, np.uint8)
    lib().or_kem_enc_batch(_ps(ps), _p(rnd), _p(pk), _p(ct), _p(ss), batch, nthreads)
    return ct.reshape(batch, -1), ss.reshape(batch, -1)


def kem_dec_batch(ps, ct, sk, nthreads=1):
    ct, sk = _u8(ct), _u8(sk)
    batch = ct.size // params(ps)["ct_bytes"]
    ss = np.zeros(batch * 32, np.uint8)
    lib().or_kem_dec_batch(_ps(ps), _p(ct), _p(sk), _p(ss), batch, nthreads)
    return ss.reshape(batch, -1)
