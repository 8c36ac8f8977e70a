"""GPU parity of the mixed-batch persistent launch (tmvp_batch_launch), its
autotuner and hipGraph capture, through the C ABI.

Every output of a mixed work list must equal, bit for bit, what the CPU oracle
gives for each job on its own (raw products mod 2^W, rounded outputs by the
closed forms pinned in tests/test_oracle.py), whatever the mix, the split of
workgroups between jobs, or the plan the autotuner picks.  Graph replays are
checked one by one against the eager result, with eager work between them.
"""
import ctypes

import numpy as np
import pytest

import oracle as O
import tmvp_amd as T

pytestmark = pytest.mark.gpu

N = 256


@pytest.fixture(scope="module")
def ctx():
    c = T.Context(0)
    yield c
    c.close()


def secret(rng, shape, ps):
    p = O.params(ps)
    v = rng.integers(-(p["mu"] // 2), p["mu"] // 2 + 1, shape)
    if p["sable"]:
        return (v % (1 << p["eq"])).astype(np.uint16)
    return v.astype(np.int16).view(np.uint16)


def public(rng, shape, bits):
    return rng.integers(0, 1 << bits, shape, dtype=np.uint16)


class Job:
    """Host inputs, device buffers, descriptor and oracle answer of one job."""

    def __init__(self, rng, op, ps, epilogue, batch, eq_bits=0):
        p = O.params(ps) if op != T.OP_POLYMUL else None
        self.op, self.ps, self.epi, self.batch = op, ps, epilogue, batch
        aux = None
        if op == T.OP_POLYMUL:
            a = public(rng, (batch, N), 16)
            s = rng.integers(-60, 61, (batch, N)).astype(np.int16).view(np.uint16)
            want = O.polymul_batch(a, s, nthreads=8) & ((1 << eq_bits) - 1)
            rows = 1
        elif op in (T.OP_MATVEC, T.OP_MATVEC_T):
            l = p["l"]
            a = public(rng, (batch, l * l * N), p["eq"])
            s = secret(rng, (batch, l * N), ps)
            raw = O.matvec_batch(l, a, s, transpose=(op == T.OP_MATVEC_T), method=1, nthreads=8)
            raw = raw.reshape(batch, l * N).astype(np.int64)
            if epilogue:
                want = ((raw + p["h1"]) % (1 << p["eq"])) >> (p["eq"] - p["ep"])
            else:
                want = raw % (1 << p["eq"])
            rows = l
        else:
            l, ep, et = p["l"], p["ep"], p["et"]
            a = public(rng, (batch, l * N), ep)
            s = secret(rng, (batch, l * N), ps)
            raw = O.innerprod_batch(l, a, s, nthreads=8).reshape(batch, N).astype(np.int64)
            if epilogue == T.IP_RAW:
                want = raw % (1 << ep)
            elif epilogue == T.IP_ENC:
                aux = rng.integers(0, 2, (batch, N), dtype=np.uint16)
                want = ((raw + p["h1"] - (aux.astype(np.int64) << (ep - 1))) % (1 << ep)) >> (ep - et)
            else:
                aux = rng.integers(0, 1 << et, (batch, N), dtype=np.uint16)
                want = ((raw + p["h2"] - (aux.astype(np.int64) << (ep - et))) % (1 << ep)) >> (ep - 1)
            rows = 1
        self.want = np.asarray(want).astype(np.uint16).reshape(batch, rows * N)
        self.da = T.DeviceBuffer.from_numpy(a)
        self.ds = T.DeviceBuffer.from_numpy(s)
        self.daux = T.DeviceBuffer.from_numpy(aux) if aux is not None else None
        self.dout = T.DeviceBuffer(batch * rows * N * 2).zero()
        self.desc = T.work(op, self.dout, self.da, self.ds, batch, ps if op != T.OP_POLYMUL else 0, epilogue,
                           self.daux, eq_bits)

    def got(self):
        return self.dout.download(np.uint16, self.want.shape)

    def free(self):
        for b in (self.da, self.ds, self.daux, self.dout):
            if b is not None:
                b.free()


def run_list(ctx, jobs, backend=T.MFMA, stream=None):
    ctx.batch_launch([j.desc for j in jobs], backend, stream=stream)
    T.device_synchronize()
    ctx.check()


def config3_jobs(rng, batch):
    """LightSaber, Saber, FireSaber instances: MatrixVectorMul rounded to p and
    InnerProd mod p of each (the mixed-parameter workload), one list."""
    jobs = []
    for ps in ("lightsaber", "saber", "firesaber"):
        jobs.append(Job(rng, T.OP_MATVEC, ps, 1, batch))
        jobs.append(Job(rng, T.OP_INNERPROD, ps, T.IP_RAW, batch))
    return jobs


@pytest.mark.parametrize("backend", [T.MFMA, T.VALU])
def test_mixed_matvec_innerprod_vs_oracle(ctx, backend):
    rng = np.random.default_rng(31)
    jobs = []
    for ps, b in (("lightsaber", 37), ("saber", 64), ("firesaber", 50)):
        jobs.append(Job(rng, T.OP_MATVEC, ps, 1, b))
        jobs.append(Job(rng, T.OP_INNERPROD, ps, T.IP_RAW, b + 3))
    run_list(ctx, jobs, backend)
    for j in jobs:
        assert np.array_equal(j.got(), j.want), (j.op, j.ps)
        j.free()


@pytest.mark.parametrize("sets", [("lightsaber", "sable2"), ("firesaber", "sable3"), ("saber", "sable1")])
def test_every_kind_in_one_launch(ctx, sets):
    """All 7 product/epilogue kinds of two parameter sets, plus a polymul:
    15 jobs of uneven sizes in one persistent launch."""
    rng = np.random.default_rng(sum(map(ord, "".join(sets))))
    jobs, b = [], 5
    for ps in sets:
        for op, epi in ((T.OP_MATVEC, 0), (T.OP_MATVEC, 1), (T.OP_MATVEC_T, 0), (T.OP_MATVEC_T, 1),
                        (T.OP_INNERPROD, T.IP_RAW), (T.OP_INNERPROD, T.IP_ENC), (T.OP_INNERPROD, T.IP_DEC)):
            jobs.append(Job(rng, op, ps, epi, b))
            b = b * 3 % 61 + 1
    jobs.append(Job(rng, T.OP_POLYMUL, None, 0, 23, eq_bits=13))
    assert len(jobs) == 15
    run_list(ctx, jobs)
    for j in jobs:
        assert np.array_equal(j.got(), j.want), (j.op, j.ps, j.epi, j.batch)
        j.free()


def test_mixed_full_size_autotune(ctx):
    """4096 instances of each Saber set: one launch equals the oracle on a
    sample and the separate single-job launches everywhere, before and after
    the autotuner installs a plan."""
    rng = np.random.default_rng(7)
    jobs = config3_jobs(rng, 4096)
    run_list(ctx, jobs)
    first = [j.got() for j in jobs]
    for j, g in zip(jobs, first):
        assert np.array_equal(g, j.want), (j.op, j.ps)
    descs = [j.desc for j in jobs]
    shape, wgs, measured = ctx.batch_plan(descs)
    assert not measured and shape in (16, 32) and wgs >= 1
    ms = ctx.batch_autotune(descs)
    assert 0.0 < ms < 50.0
    shape, wgs, measured = ctx.batch_plan(descs)  # the plan records the tile it was measured with
    assert measured and shape in (16, 32) and wgs >= 1
    for j in jobs:
        j.dout.zero()
    run_list(ctx, jobs)
    for j, g in zip(jobs, first):
        assert np.array_equal(j.got(), g), (j.op, j.ps)
        j.free()


@pytest.mark.parametrize("tile", [T.TILE_16, T.TILE_32], ids=["tile16", "tile32"])
def test_mixed_launch_each_tile(ctx, tile):
    """Every kind in one launch with the descriptors naming the MFMA tile: the
    16x16x64 and the 32x32x32 multi-job kernels both equal the oracle; the
    plan query reports the tile, and descriptors naming two tiles are refused."""
    rng = np.random.default_rng(70 + tile)
    jobs = config3_jobs(rng, 97)
    for j in jobs:
        j.desc.flags |= tile
    run_list(ctx, jobs)
    for j in jobs:
        assert np.array_equal(j.got(), j.want), (j.op, j.ps, tile)
    shape, _, _ = ctx.batch_plan([j.desc for j in jobs])
    assert shape == (32 if tile == T.TILE_32 else 16)
    jobs[0].desc.flags ^= T.TILE_16 | T.TILE_32
    with pytest.raises(T.TmvpError) as e:
        ctx.batch_launch([j.desc for j in jobs])
    assert e.value.code == T.EARG
    for j in jobs:
        j.free()


def test_graph_replay_matches_eager(ctx):
    """A captured mixed launch plus a KEM encapsulation, replayed three times
    with eager work (clearing the outputs, an unrelated launch) in between."""
    rng = np.random.default_rng(8)
    jobs = config3_jobs(rng, 257)
    st = T.Stream()
    kb = 300
    pk, sk = O.kem_keypair_batch("saber", rng.integers(0, 256, (kb, 96), dtype=np.uint8), nthreads=8)
    rnd = rng.integers(0, 256, (kb, 32), dtype=np.uint8)
    want_ct, want_ss = O.kem_enc_batch("saber", rnd, pk, nthreads=8)
    dpk, drnd, dsk = T.DeviceBuffer.from_numpy(pk), T.DeviceBuffer.from_numpy(rnd), T.DeviceBuffer.from_numpy(sk)
    dct, dss, dss2, dfail = T.DeviceBuffer(kb * 1088), T.DeviceBuffer(kb * 32), T.DeviceBuffer(kb * 32), T.DeviceBuffer(kb)
    ctx.kem_reserve("saber", kb, stream=st)

    def body():  # the decapsulation forks to its side stream and joins back inside the capture
        ctx.batch_launch([j.desc for j in jobs], stream=st)
        ctx.kem_encaps(dct, dss, dpk, drnd, kb, "saber", T.MFMA, stream=st)
        ctx.kem_decaps(dss2, dct, dsk, kb, "saber", T.MFMA, fail=dfail, stream=st)

    g = ctx.capture(st, body)
    other = Job(rng, T.OP_MATVEC, "saber", 0, 99)
    try:
        for rep in range(3):
            for j in jobs:
                j.dout.zero()
            dct.zero()
            dss.zero()
            dss2.zero()
            dfail.zero()
            T.device_synchronize()
            g.launch(st)
            st.synchronize()
            ctx.check(st)
            for j in jobs:
                assert np.array_equal(j.got(), j.want), (rep, j.op, j.ps)
            assert np.array_equal(dct.download(np.uint8, (kb, -1)), want_ct), rep
            assert np.array_equal(dss.download(np.uint8, (kb, -1)), want_ss), rep
            assert np.array_equal(dss2.download(np.uint8, (kb, -1)), want_ss), rep
            assert not dfail.download(np.uint8).any(), rep
            run_list(ctx, [other])  # eager work between replays
            assert np.array_equal(other.got(), other.want)
    finally:
        g.destroy()
        st.destroy()
    for b in (dpk, drnd, dsk, dct, dss, dss2, dfail):
        b.free()
    for j in jobs + [other]:
        j.free()


def test_graph_keeps_its_workspace_when_the_stream_grows(ctx):
    """A graph holding a KEM capture stays correct after the capture stream's
    workspace grows for a larger eager call, and after more streams than the
    context keeps workspaces for have evicted it: the graph's buffer is kept
    alive and not reused."""
    rng = np.random.default_rng(81)
    st = T.Stream()
    kb, big = 200, 3000
    pk, _ = O.kem_keypair_batch("saber", rng.integers(0, 256, (big, 96), dtype=np.uint8), nthreads=8)
    rnd = rng.integers(0, 256, (big, 32), dtype=np.uint8)
    want_ct, want_ss = O.kem_enc_batch("saber", rnd[:kb], pk[:kb], nthreads=8)
    dpk, drnd = T.DeviceBuffer.from_numpy(pk), T.DeviceBuffer.from_numpy(rnd)
    dct, dss = T.DeviceBuffer(kb * 1088), T.DeviceBuffer(kb * 32)
    ect, ess = T.DeviceBuffer(big * 1088), T.DeviceBuffer(big * 32)
    ctx.kem_reserve("saber", kb, stream=st)
    g = ctx.capture(st, lambda: ctx.kem_encaps(dct, dss, dpk, drnd, kb, "saber", T.MFMA, stream=st))
    others = []
    try:
        for rep in range(3):
            if rep == 1:  # grow the capture stream's workspace
                ctx.kem_encaps(ect, ess, dpk, drnd, big, "saber", T.MFMA, stream=st)
                st.synchronize()
            if rep == 2:  # evict it: 17 more streams each get a workspace
                others = [T.Stream() for _ in range(17)]
                for o in others:
                    ctx.kem_encaps(ect, ess, dpk, drnd, 64, "saber", T.MFMA, stream=o)
                T.device_synchronize()
            dct.zero()
            dss.zero()
            T.device_synchronize()
            g.launch(st)
            st.synchronize()
            ctx.check(st)
            assert np.array_equal(dct.download(np.uint8, (kb, -1)), want_ct), rep
            assert np.array_equal(dss.download(np.uint8, (kb, -1)), want_ss), rep
    finally:
        g.destroy()
        for o in others:
            o.destroy()
        st.destroy()
    for b in (dpk, drnd, dct, dss, ect, ess):
        b.free()


def test_capture_refuses_workspace_growth(ctx):
    st = T.Stream()
    kb = 5000
    bufs = [T.DeviceBuffer(kb * n) for n in (1312, 32, 1472, 32)]
    lib = T.load_library()
    assert lib.tmvp_graph_begin(ctx.handle, st.handle) == 0
    rc = lib.tmvp_kem_encaps(ctx.handle, T.FIRESABER, T.MFMA, bufs[2].ptr, bufs[3].ptr, bufs[0].ptr, bufs[1].ptr,
                             kb, st.handle)
    g = ctypes.c_void_p()
    lib.tmvp_graph_end(ctx.handle, st.handle, ctypes.byref(g))
    if g.value:
        lib.tmvp_graph_destroy(g)
    assert rc == -4
    st.destroy()
    for b in bufs:
        b.free()


def test_batch_argument_errors(ctx):
    rng = np.random.default_rng(1)
    j = Job(rng, T.OP_MATVEC, "saber", 1, 3)
    lib = T.load_library()

    def launch(descs, backend=T.MFMA):
        arr = (T.WorkDesc * max(1, len(descs)))(*descs)
        return lib.tmvp_batch_launch(ctx.handle, backend, ctypes.cast(arr, ctypes.c_void_p), len(descs), None)

    assert launch([j.desc] * 17) == -1                       # more than TMVP_MAX_WORK
    bad = T.WorkDesc.from_buffer_copy(j.desc)
    bad.op = 9
    assert launch([j.desc, bad]) == -1
    bad = T.WorkDesc.from_buffer_copy(j.desc)
    bad.flags = T.ACCUMULATE                                  # rounded output cannot accumulate
    assert launch([bad]) == -1
    poly = T.work(T.OP_POLYMUL, j.dout, j.da, j.ds, 3, eq_bits=7)
    assert launch([poly]) == -4                               # MFMA needs 8 <= eq_bits <= 14
    assert launch([poly], T.VALU) == 0
    assert launch([]) == 0
    assert lib.tmvp_batch_launch(ctx.handle, 5, None, 0, None) == -1
    acc = T.work(T.OP_MATVEC, j.dout, j.da, j.ds, 3, "saber", 0, flags=T.ACCUMULATE)
    arr = (T.WorkDesc * 1)(acc)
    assert lib.tmvp_batch_autotune(ctx.handle, ctypes.cast(arr, ctypes.c_void_p), 1, None, None) == -1
    assert lib.tmvp_graph_begin(ctx.handle, None) == -1
    T.device_synchronize()
    j.free()
