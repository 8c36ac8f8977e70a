#!/usr/bin/env python3
"""Benchmark: Saber MatrixVectorMul throughput at B = 65536 on one MI355X.

One step = one launch of the batched Saber (L=3) MatrixVectorMul with the
fused q->p rounding over 65536 independent instances whose inputs (A: 3x3
uniform 13-bit polynomials, s: 3 CBD(mu=8) secrets) are already resident in
HBM; the output is the rounded b' = ((A s + h1) mod q) >> 3 per instance.  The
headline `value` is the MFMA form's MatVec/s; the plain VALU form is timed
beside it on the same buffers (`valu`), and the two outputs are compared bit
for bit before timing.

Timing uses HIP events recorded on the stream the kernels are launched on.
The K timed launches run twice: right after the W warm-up launches
(`ms_per_step_cold`, the device still in its start-up clock transient) and
after `settle_launches` more back-to-back launches (`ms_per_step`, the settled
device, which `value` reports).  `mfma_util_of_published` is the executed
int8 rate (both limbs) against MI355X's published 5.0 POPS; `roofline.measured`
holds this device's probed ceilings (HBM streaming with this traffic shape,
both int8 MFMA shapes) with the shape each was best at.

`polymul` is config 2: 65536 raw products, MFMA against VALU, bit for bit.
`batch1` is config 1's shape on the GPU: one MatrixVectorMul per call.
`cpu_baseline` times the CPU oracle (a C restatement of the reference's
Toom-4 and schoolbook paths) on bounded samples: the headline shape at one
thread and at the process's host-core share, raw products, and one
instance at a time (config 1).

`mixed` reports the mixed-parameter workload: 4096 LightSaber, 4096 Saber and
4096 FireSaber instances, each a MatrixVectorMul rounded to p plus an
InnerProd mod p, as one persistent launch (default and autotuned plan, and
replayed from a hipGraph) beside the six separate launches.

`serving` drives the serving loop with Poisson arrivals at a fixed rate:
half Saber, half Sable type 2, half encapsulations, half decapsulations,
request sizes uniform in [1, 4096], keys drawn from 100k device-resident keys
per set; it reports p50/p99 request latency (submit to results in host
memory) and instances/s.

`kem` reports the full batched Saber KEM at 16384 instances (key generation,
encapsulation, decapsulation with all sampling, hashing, codecs and products on
the GPU; seeds and keys resident in HBM): ops/s per step and p50/p99 of the
GPU time of each batched call, after a round-trip check of every instance,
with its Keccak-f[1600] rate against the VALU-only permutation probe.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--settle S] [--no-kem] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "asfand-tmvp_amd"))
# Eight hardware queues for this process (HIP's default is four), read when
# the runtime initialises: the serving leg's slot streams then each get their
# own queue.  With four, its p50 at 6000 req/s moved between 0.35 and 0.55 ms
# with where the runtime happened to place the streams.
os.environ["GPU_MAX_HW_QUEUES"] = "8"

import tmvp_amd as T  # noqa: E402

N = 256
L = 3
BATCH = 65536
# Algorithmic HBM bytes per instance: A (9 x 256 x 2) + s (3 x 256 x 2) + out (3 x 256 x 2).
BYTES_PER_INSTANCE = 9 * N * 2 + 3 * N * 2 + 3 * N * 2
MACS_PER_INSTANCE = 9 * N * N
# Packed 16-bit multiply-adds per second of the VALU at 2.4 GHz (256 CU x 4 SIMD x 16 lanes x 2).
VALU_PEAK_MACS = 256 * 4 * 16 * 2.4e9 * 2
HBM_PEAK_GBS = 8000.0
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "mfma_matvec_saber_b65536_hbm_traffic.md")


def measured_traffic():
    """HBM bytes per launch of the MFMA kernel, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (scaled by the copy-kernel calibration)."""
    try:
        for line in open(TRAFFIC_PROFILE):
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


def synthetic_inputs(batch, seed=1):
    rng = np.random.default_rng(seed)
    A = rng.integers(0, 8192, (batch, L * L * N), dtype=np.uint16)
    # CBD(mu=8): popcount of 4 bits minus popcount of 4 bits, as int16 two's complement
    bits = rng.integers(0, 256, (batch, L * N), dtype=np.uint8)
    pc = np.unpackbits(bits[..., None], axis=-1).astype(np.int16)
    s = (pc[..., :4].sum(-1) - pc[..., 4:].sum(-1)).astype(np.int16).view(np.uint16)
    return A, s


# MI355X published dense peaks (the roofline's `peak`); the probes below are
# what this device reaches with this workload's shapes.
MFMA_I8_PUBLISHED_TOPS = 5000.0


def peaks_lib():
    import ctypes
    path = os.path.join(ROOT, "asfand-tmvp_amd", "libtmvp_peaks.so")
    if not os.path.exists(path):
        return None
    lib = ctypes.CDLL(path)
    lib.tmvp_peak_keccak.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    return lib


def measured_peaks():
    """HBM streaming rate with this workload's traffic shape (at 1-4
    workgroups of four waves per CU) and the int8 matrix-core rate of both
    gfx950 int8 MFMA shapes (at 1, 2 and 4 waves per SIMD), measured on this
    device (libtmvp_peaks.so); the best of each is the measured ceiling."""
    import ctypes
    lib = peaks_lib()
    if lib is None:
        return None
    us, gbs, tops = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
    stream = {}
    for w in (1, 2, 3, 4):
        if lib.tmvp_peak_stream_at(ctypes.c_int64(BATCH), 20, w, ctypes.byref(us), ctypes.byref(gbs)):
            return None
        stream[f"{w}_wg_per_cu"] = {"us": us.value, "GBs": gbs.value}
    mfma = {}
    for shape, name in ((0, "16x16x64"), (1, "32x32x32")):
        for wps in (1, 2, 4):
            if lib.tmvp_peak_mfma_i8_shape(5, shape, wps, ctypes.byref(tops)):
                return None
            mfma[f"{name}_{wps}w"] = tops.value
    best_stream = min(stream.items(), key=lambda kv: kv[1]["us"])
    best_mfma = max(mfma.items(), key=lambda kv: kv[1])
    return {"hbm_stream_GBs": best_stream[1]["GBs"], "hbm_stream_us": best_stream[1]["us"],
            "stream_shape": f"per instance 6144 B read + 1536 B written, one wave per instance, B=65536, "
                            f"best at {best_stream[0]}",
            "stream_by_shape": stream,
            "mfma_i8_tops": best_mfma[1], "mfma_i8_shape": best_mfma[0], "mfma_i8_by_shape": mfma}


def keccak_peak():
    """Keccak-f[1600] permutations/s of the KEM's lane-pair round code on the
    VALU alone (state in registers), best over 1-8 waves per SIMD."""
    import ctypes
    lib = peaks_lib()
    if lib is None:
        return None
    rate = ctypes.c_double()
    by = {}
    for wps in (1, 2, 4, 8):
        for u in (1, 2):
            if lib.tmvp_peak_keccak(5, wps, u, ctypes.byref(rate)):
                return None
            by[f"{wps}w_u{u}"] = rate.value
    best = max(by.items(), key=lambda kv: kv[1])
    return {"perms_per_s": best[1], "shape": best[0], "by_shape": by}


def keccak_perms(p):
    """Keccak-f[1600] calls per instance of each KEM operation, from the
    sponge lengths (SHAKE-128 rate 168, SHA3-256 136, SHA3-512 72 bytes)."""
    shake = lambda n: -(-n // 168)
    h = lambda n: n // 136 + 1
    g = lambda n: n // 72 + 1
    araw, sraw = p["l"] * p["l"] * p["eq"] * 32, p["l"] * p["mu"] * 32
    return {"keypair": shake(32) + shake(araw) + shake(sraw) + h(p["pk_bytes"]),
            "encaps": h(32) + h(p["pk_bytes"]) + g(64) + shake(araw) + shake(sraw) + h(p["ct_bytes"]) + h(64),
            "decaps": g(64) + shake(araw) + shake(sraw) + h(p["ct_bytes"]) + h(64)}


def time_launches(fn, stream, steps, warmup):
    for _ in range(warmup):
        fn()
    stream.synchronize()
    T.device_synchronize()
    start, stop = T.Event(), T.Event()
    start.record(stream)
    for _ in range(steps):
        fn()
    stop.record(stream)
    stop.synchronize()
    T.device_synchronize()
    ms = start.elapsed_ms(stop)
    start.destroy()
    stop.destroy()
    return ms


# A one-GPU process is given a 16-core share of its host, whatever the
# machine's core count: the thread pools use at most that many.
CPU_SHARE = 16


def host_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_rate(fn, batch, seconds):
    """Instances per second of fn() (one call = `batch` instances) over about
    `seconds` of wall time, after one untimed call."""
    fn()
    done, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        fn()
        done += batch
    dt = time.perf_counter() - t0
    return done / dt, done, dt


def cpu_baseline():
    """Time the CPU oracle (the C restatement of the reference's schoolbook
    and Toom-4 paths) on host cores over bounded samples of the same
    workloads: Saber L=3 MatrixVectorMul (the headline shape) and raw products
    (config 2) at one thread and at the whole pool, and one MatrixVectorMul
    at a time on one thread (config 1, batch 1)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # test infrastructure: used only as the timed CPU baseline

    ncpu = host_cores()
    pool = max(1, min(ncpu, CPU_SHARE))
    A, s = synthetic_inputs(4096, seed=2)
    legs = {}
    samples = []
    for method, mname, secs in ((1, "toom4", 6.0), (0, "schoolbook", 3.0)):
        for threads, tname, n in ((pool, "pool", 4096), (1, "1thread", 256)):
            sec = secs if threads == pool else secs / 2
            rate, done, dt = cpu_rate(lambda: O.matvec_batch(L, A[:n], s[:n], method=method, nthreads=threads),
                                      n, sec)
            legs[f"matvec_{mname}_{tname}_per_s"] = rate
            samples.append(f"{done} MatVec {mname} {threads}t {dt:.1f}s")
    a = A.reshape(-1, N)[:8192]
    sp = s.reshape(-1, N)[:8192]
    for threads, tname, n in ((pool, "pool", 8192), (1, "1thread", 1024)):
        rate, done, dt = cpu_rate(lambda: O.polymul_batch(a[:n], sp[:n], method=1, nthreads=threads), n, 2.0)
        legs[f"polymul_toom4_{tname}_per_s"] = rate
        samples.append(f"{done} polymul toom4 {threads}t {dt:.1f}s")
    for method, mname in ((0, "schoolbook"), (1, "toom4")):
        O.bench_matvec1(L, A[0], s[0], method, 200)
        legs[f"batch1_matvec_{mname}_ns"] = O.bench_matvec1(L, A[0], s[0], method, 20000 if method else 10000)
    return {"value": legs["matvec_toom4_pool_per_s"], "unit": "MatVec/s", "cores": pool, "kind": "port",
            "host_cores": ncpu,
            "sample": f"Saber L=3 MatrixVectorMul, Toom-4 over Karatsuba, {pool} threads (the GPU's {CPU_SHARE}-core "
                      f"host share; the host has {ncpu}); legs: " + "; ".join(samples),
            "legs": legs}


def percentile(xs, q):
    xs = sorted(xs)
    k = min(len(xs) - 1, max(0, int(np.ceil(q / 100.0 * len(xs))) - 1))
    return xs[k]


def kem_bench(ctx, stream, batch=16384, reps=100, ps="saber", cpu_seconds=4.0, with_cpu=True):
    """Batched KEM: keypair, encaps and decaps of `batch` instances per call."""
    p = T.get_params(ps)
    rng = np.random.default_rng(3)
    seeds = T.DeviceBuffer.from_numpy(rng.integers(0, 256, (batch, 96), dtype=np.uint8))
    rnd = T.DeviceBuffer.from_numpy(rng.integers(0, 256, (batch, 32), dtype=np.uint8))
    pk, sk = T.DeviceBuffer(batch * p["pk_bytes"]), T.DeviceBuffer(batch * p["sk_bytes"])
    ct, ss, ss2 = T.DeviceBuffer(batch * p["ct_bytes"]), T.DeviceBuffer(batch * 32), T.DeviceBuffer(batch * 32)
    fail = T.DeviceBuffer(batch)

    def keypair():
        ctx.kem_keypair(pk, sk, seeds, batch, ps, T.MFMA, stream=stream)

    def encaps():
        ctx.kem_encaps(ct, ss, pk, rnd, batch, ps, T.MFMA, stream=stream)

    def decaps():
        ctx.kem_decaps(ss2, ct, sk, batch, ps, T.MFMA, fail=fail, stream=stream)

    keypair(); encaps(); decaps()
    stream.synchronize()
    ctx.check(stream)
    ok = (not fail.download(np.uint8).any()) and np.array_equal(ss.download(np.uint8), ss2.download(np.uint8))
    if not ok:
        print("[bench] ERROR: KEM round trip failed", file=sys.stderr)
        sys.exit(1)

    ev = [T.Event() for _ in range(4)]
    t_kp, t_enc, t_dec, t_pair = [], [], [], []
    for _ in range(10):
        keypair(); encaps(); decaps()
    stream.synchronize()
    for _ in range(reps):
        ev[0].record(stream); keypair()
        ev[1].record(stream); encaps()
        ev[2].record(stream); decaps()
        ev[3].record(stream)
        ev[3].synchronize()
        t_kp.append(ev[0].elapsed_ms(ev[1]))
        t_enc.append(ev[1].elapsed_ms(ev[2]))
        t_dec.append(ev[2].elapsed_ms(ev[3]))
        t_pair.append(ev[1].elapsed_ms(ev[3]))
    for e in ev:
        e.destroy()
    ctx.check(stream)
    perms = keccak_perms(p)
    kf = batch * (perms["keypair"] + perms["encaps"] + perms["decaps"]) / (
        (np.mean(t_kp) + np.mean(t_enc) + np.mean(t_dec)) / 1e3)
    out = {
        "workload": f"{ps} KEM keypair/encaps/decaps, {batch} instances per call",
        "batch": batch, "reps": reps,
        "keccak_f_per_instance": perms,
        "keccak_f_per_s": kf,
        "keypair_per_s": batch / (np.mean(t_kp) / 1e3),
        "encaps_per_s": batch / (np.mean(t_enc) / 1e3),
        "decaps_per_s": batch / (np.mean(t_dec) / 1e3),
        "encaps_decaps_per_s": batch / (np.mean(t_pair) / 1e3),
        "keypair_ms": {"p50": percentile(t_kp, 50), "p99": percentile(t_kp, 99)},
        "encaps_ms": {"p50": percentile(t_enc, 50), "p99": percentile(t_enc, 99)},
        "decaps_ms": {"p50": percentile(t_dec, 50), "p99": percentile(t_dec, 99)},
        "encaps_decaps_ms": {"p50": percentile(t_pair, 50), "p99": percentile(t_pair, 99)},
        "round_trip_ok": ok,
    }
    for b in (seeds, rnd, pk, sk, ct, ss, ss2, fail):
        b.free()
    if with_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O  # test infrastructure: used only as the timed CPU baseline

        cores = max(1, min(host_cores(), CPU_SHARE))
        n = 1024
        seeds_h = np.random.default_rng(4).integers(0, 256, (n, 96), dtype=np.uint8)
        rnd_h = np.random.default_rng(5).integers(0, 256, (n, 32), dtype=np.uint8)
        pk_h, sk_h = O.kem_keypair_batch(ps, seeds_h, nthreads=cores)
        done, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < cpu_seconds:
            ct_h, _ = O.kem_enc_batch(ps, rnd_h, pk_h, nthreads=cores)
            O.kem_dec_batch(ps, ct_h, sk_h, nthreads=cores)
            done += n
        out["cpu_oracle_encaps_decaps_per_s"] = done / (time.perf_counter() - t0)
        out["cpu_oracle_cores"] = cores
    return out


# Algorithmic bytes per instance of the mixed workload: A + s + b' + b + v.
MIXED_BYTES = {2: 5632, 3: 9728, 4: 14848}


def mixed_bench(ctx, stream, batch=4096, steps=50):
    rng = np.random.default_rng(6)
    bufs, descs, singles = [], [], []
    total_bytes = 0
    for ps, l, mu in (("lightsaber", 2, 10), ("saber", 3, 8), ("firesaber", 4, 6)):
        A = rng.integers(0, 8192, (batch, l * l * N), dtype=np.uint16)
        bits = rng.integers(0, 1 << mu, (batch, l * N))
        half = mu // 2
        pc = lambda x: np.unpackbits(x.astype(">u2").view(np.uint8).reshape(*x.shape, 2), axis=-1).sum(-1)
        s = (pc(bits & ((1 << half) - 1)) - pc(bits >> half)).astype(np.int16).view(np.uint16)
        b = rng.integers(0, 1024, (batch, l * N), dtype=np.uint16)
        dA, ds_, db = T.DeviceBuffer.from_numpy(A), T.DeviceBuffer.from_numpy(s), T.DeviceBuffer.from_numpy(b)
        dbp, dv = T.DeviceBuffer(batch * l * N * 2), T.DeviceBuffer(batch * N * 2)
        bufs += [dA, ds_, db, dbp, dv]
        descs.append(T.work(T.OP_MATVEC, dbp, dA, ds_, batch, ps, 1))
        descs.append(T.work(T.OP_INNERPROD, dv, db, ds_, batch, ps, T.IP_RAW))
        singles.append((dbp, dA, ds_, db, dv, ps))
        total_bytes += MIXED_BYTES[l] * batch

    def separate():
        for dbp, dA, ds_, db, dv, ps in singles:
            ctx.matvec(dbp, dA, ds_, batch, ps, T.MFMA, round_to_p=True, stream=stream)
            ctx.innerprod(dv, db, ds_, batch, ps, T.MFMA, T.IP_RAW, stream=stream)

    def fused():
        ctx.batch_launch(descs, T.MFMA, stream=stream)

    out = {"workload": f"{batch} each LightSaber/Saber/FireSaber: MatVec round-to-p + InnerProd mod p",
           "instances": 3 * batch, "algorithmic_bytes": total_bytes}
    ms_sep = time_launches(separate, stream, steps, 5) / steps
    ms_def = time_launches(fused, stream, steps, 5) / steps
    fused()
    stream.synchronize()
    ref = [x.download(np.uint16) for x in bufs[3::5] + bufs[4::5]]
    separate()
    stream.synchronize()
    same = all(np.array_equal(a, x.download(np.uint16)) for a, x in zip(ref, bufs[3::5] + bufs[4::5]))
    tuned = ctx.batch_autotune(descs, stream=stream)
    tile, wgs, _ = ctx.batch_plan(descs)
    ms_tuned = time_launches(fused, stream, steps, 5) / steps
    g = ctx.capture(stream, fused)
    ms_graph = time_launches(lambda: g.launch(stream), stream, steps, 5) / steps
    g.destroy()
    ctx.check(stream)
    best = min(ms_def, ms_tuned, ms_graph)
    out.update({
        "separate_launches_ms": ms_sep, "persistent_default_ms": ms_def, "persistent_autotuned_ms": ms_tuned,
        "autotune_best_ms": tuned, "autotuned_plan": {"mfma_tile": tile, "wgs_per_cu": wgs},
        "graph_replay_ms": ms_graph,
        "instances_per_s": 3 * batch / (best / 1e3),
        "achieved_GBs": total_bytes / (best / 1e3) / 1e9,
        "hbm_frac": total_bytes / (best / 1e3) / 1e9 / HBM_PEAK_GBS,
        "outputs_equal_separate": bool(same),
    })
    for x in bufs:
        x.free()
    if not same:
        print("[bench] ERROR: mixed launch differs from separate launches", file=sys.stderr)
        sys.exit(1)
    return out


def graph_chain_bench(ctx, stream, calls=32, steps=50):
    """Where a graph pays: a chain of short launches.  calls single-instance
    Saber MatrixVectorMul launches (config 1's shape, each on its own
    instance) back to back, timed eager and as one replayed graph."""
    A, s = synthetic_inputs(calls, seed=13)
    dA, ds_ = T.DeviceBuffer.from_numpy(A), T.DeviceBuffer.from_numpy(s)
    o1, o2 = T.DeviceBuffer(calls * L * N * 2), T.DeviceBuffer(calls * L * N * 2)
    ea, es = L * L * N * 2, L * N * 2  # bytes per instance of A and s

    def chain(out):
        base_o, base_a, base_s = out.ptr.value, dA.ptr.value, ds_.ptr.value
        for i in range(calls):
            ctx.matvec(base_o + i * es, base_a + i * ea, base_s + i * es, 1, T.SABER, T.MFMA, round_to_p=True,
                       stream=stream)

    chain(o1)
    stream.synchronize()
    g = ctx.capture(stream, lambda: chain(o2))
    ms_eager = time_launches(lambda: chain(o1), stream, steps, 5) / steps
    ms_graph = time_launches(lambda: g.launch(stream), stream, steps, 5) / steps
    stream.synchronize()
    same = np.array_equal(o1.download(np.uint16), o2.download(np.uint16))
    g.destroy()
    ctx.check(stream)
    for b in (dA, ds_, o1, o2):
        b.free()
    if not same:
        print("[bench] ERROR: graph replay differs from the eager launches", file=sys.stderr)
        sys.exit(1)
    return {"workload": f"{calls} single-instance Saber MatrixVectorMul launches back to back",
            "eager_ms": ms_eager, "graph_replay_ms": ms_graph, "outputs_equal": bool(same)}


def serving_bench(ctx, qps=2000.0, seconds=3.0, nkeys=100_000, max_count=4096, seed=9):
    sets = ("saber", "sable2")
    rng = np.random.default_rng(seed)
    srv = T.Server(ctx, max_batch=16384)  # the server's default of four slots
    pools, keys = {}, {}
    try:
        t0 = time.perf_counter()
        for ps in sets:
            p = T.get_params(ps)
            pk, sk = T.run_kem_keypair(ctx, rng.integers(0, 256, (nkeys, 96), dtype=np.uint8), ps, T.MFMA)
            srv.add_keys(ps, pk, sk)
            npool = 16384
            ids = rng.integers(0, nkeys, npool).astype(np.int32)
            rnd = T.pinned_array((npool, 32))
            rnd[:] = rng.integers(0, 256, (npool, 32), dtype=np.uint8)
            ct = T.pinned_array((npool, p["ct_bytes"]))
            ct[:] = T.run_kem_encaps(ctx, pk[ids], rng.integers(0, 256, (npool, 32), dtype=np.uint8), ps, T.MFMA)[0]
            pools[ps] = (ids, rnd, ct)
            keys[ps] = (pk, sk)
        setup_s = time.perf_counter() - t0
        n_req = int(qps * seconds)
        gaps = rng.exponential(1.0 / qps, n_req)
        arrivals = np.cumsum(gaps)
        plan = []
        for k in range(n_req):
            ps = sets[int(rng.integers(0, 2))]
            op = int(rng.integers(0, 2))
            n = int(rng.integers(1, max_count + 1))
            off = int(rng.integers(0, 16384 - n + 1))
            ids, rnd, ct = pools[ps]
            plan.append((op, ps, ids[off:off + n], (rnd if op == T.SRV_ENCAPS else ct)[off:off + n]))
        # outputs land in a pinned ring; after the run the last requests whose
        # region was not reused are checked against the plain batched calls
        ring = T.pinned_array((256 << 20,))
        ring_off, ring_gen = 0, 0
        placed = {}

        def outputs(op, ps, n, tag=None):
            nonlocal ring_off, ring_gen
            ctb = T.get_params(ps)["ct_bytes"] if op == T.SRV_ENCAPS else 1
            need = n * (ctb + 32)
            if ring_off + need > ring.size:
                ring_off, ring_gen = 0, ring_gen + 1
            a = ring[ring_off:ring_off + n * ctb].reshape(n, ctb)
            b = ring[ring_off + n * ctb:ring_off + need].reshape(n, 32)
            out = (a, b) if op == T.SRV_ENCAPS else (b, a.reshape(n))
            if tag is not None:
                placed[tag] = (ring_gen, ring_off, out)
            ring_off += need
            return out

        # warm-up: one small and one full-size request of every (set, operation),
        # through the same input and output paths as the timed requests (the
        # first copies into a caller's pinned buffer carry a one-time cost)
        for ps in sets:
            ids, rnd, ct = pools[ps]
            for op in (T.SRV_ENCAPS, T.SRV_DECAPS):
                for n in (1, max_count):
                    srv.submit(op, ps, ids[:n], (rnd if op == T.SRV_ENCAPS else ct)[:n], tag=1 << 40,
                               out=outputs(op, ps, n))
        srv.drain()
        while srv.poll(256):
            pass
        done = []
        start = time.perf_counter()
        next_poll = 0.0
        for k, (op, ps, ids, data) in enumerate(plan):
            while True:
                now = time.perf_counter() - start
                if now >= arrivals[k]:
                    break
                if now >= next_poll:  # collect completions at most every 0.5 ms
                    done += srv.poll(256)
                    next_poll = now + 5e-4
            srv.submit(op, ps, ids, data, tag=k, out=outputs(op, ps, len(ids), tag=k))
        submit_lag_ms = (time.perf_counter() - start - arrivals[-1]) * 1e3
        while len(done) < n_req:
            done += srv.poll(256, timeout_us=100000)
        st = srv.stats()
        st["batches"] -= 8
        st["instances_done"] -= 4 * (1 + max_count)
    finally:
        srv.close()
    # the last (up to) 16 requests whose outputs are still in the ring, against
    # plain batched calls with the same keys and inputs
    intact = [k for k in range(n_req - 1, -1, -1)
              if placed[k][0] == ring_gen or (placed[k][0] == ring_gen - 1 and ring_off <= placed[k][1])][:16]
    checked_ok = True
    for k in intact:
        op, ps, ids, data = plan[k]
        pk, sk = keys[ps]
        got = placed[k][2]
        if op == T.SRV_ENCAPS:
            want = T.run_kem_encaps(ctx, pk[ids], np.asarray(data), ps, T.MFMA)
        else:
            want = T.run_kem_decaps(ctx, np.asarray(data), sk[ids], ps, T.MFMA)
        checked_ok = checked_ok and all(np.array_equal(g, w) for g, w in zip(got, want))
    if not checked_ok:
        print("[bench] ERROR: serving outputs differ from the batched calls", file=sys.stderr)
        sys.exit(1)
    bad = sum(1 for c in done if c["status"] != 0)
    lat = sorted((c["done_us"] - c["submit_us"]) / 1e3 for c in done)
    queue = sorted((c["start_us"] - c["submit_us"]) / 1e3 for c in done)
    service = sorted((c["done_us"] - c["start_us"]) / 1e3 for c in done)
    span_s = (max(c["done_us"] for c in done) - min(c["submit_us"] for c in done)) / 1e6
    inst = sum(len(x[2]) for x in plan)
    by_op = {}
    for name, op in (("encaps", T.SRV_ENCAPS), ("decaps", T.SRV_DECAPS)):
        l = sorted((c["done_us"] - c["submit_us"]) / 1e3 for c in done if plan[c["tag"]][0] == op)
        by_op[name] = {"p50_ms": percentile(l, 50), "p99_ms": percentile(l, 99)}
    return {
        "workload": f"Poisson {qps:.0f} req/s for {seconds:.0f} s; 50% Saber / 50% Sable-2, 50% encaps / 50% decaps, "
                    f"request size U[1,{max_count}], {nkeys} resident keys per set",
        "requests": n_req, "instances": inst, "errors": bad,
        "offered_requests_per_s": qps, "submit_lag_ms": submit_lag_ms,
        "p50_ms": percentile(lat, 50), "p99_ms": percentile(lat, 99), "max_ms": lat[-1],
        "p90_ms": percentile(lat, 90), "p999_ms": percentile(lat, 99.9),
        "queue_ms": {"p50": percentile(queue, 50), "p99": percentile(queue, 99)},
        "service_ms": {"p50": percentile(service, 50), "p99": percentile(service, 99)},
        "instances_per_s": inst / span_s, "requests_per_s": n_req / span_s,
        "saturated": n_req / span_s < 0.95 * qps,
        "by_op": by_op, "batches": st["batches"], "mean_batch_instances": st["instances_done"] / max(1, st["batches"]),
        "slots": st["slots"], "max_concurrent_batches": st["max_concurrent"], "setup_s": setup_s,
        "hbm_budget": st["hbm_budget"], "store_bytes": st["store_bytes"], "arena_bytes": st["arena_bytes"],
        "max_inflight_bytes": st["max_inflight_bytes"],
        "outputs_checked": len(intact), "outputs_equal_batched": bool(checked_ok),
    }


def polymul_bench(ctx, stream, steps, batch=65536):
    """Raw batched negacyclic products (config 2): B = 65536 products of a
    uniform 13-bit a by a CBD(mu=8) s, out mod 2^13, MFMA and VALU forms on
    the same buffers, outputs compared bit for bit."""
    rng = np.random.default_rng(12)
    a = rng.integers(0, 8192, (batch, N), dtype=np.uint16)
    _, s3 = synthetic_inputs(batch // 3 + 1, seed=13)
    s = s3.reshape(-1, N)[:batch]
    da, ds_ = T.DeviceBuffer.from_numpy(a), T.DeviceBuffer.from_numpy(s)
    om, ov = T.DeviceBuffer(batch * N * 2), T.DeviceBuffer(batch * N * 2)
    run_m = lambda: ctx.polymul(om, da, ds_, batch, 13, T.MFMA, stream=stream)
    run_v = lambda: ctx.polymul(ov, da, ds_, batch, 13, T.VALU, stream=stream)
    run_m(); run_v()
    stream.synchronize()
    ctx.check(stream)
    same = bool(np.array_equal(om.download(np.uint16), ov.download(np.uint16)))
    if not same:
        print("[bench] ERROR: polymul MFMA and VALU outputs differ", file=sys.stderr)
        sys.exit(1)
    ms_m = time_launches(run_m, stream, steps, 20) / steps
    ms_v = time_launches(run_v, stream, max(5, steps // 4), 3) / max(5, steps // 4)
    for b in (da, ds_, om, ov):
        b.free()
    algo = 1536 * batch
    return {"workload": f"{batch} negacyclic products a*s mod (x^256+1, 2^13), a uniform 13-bit, s CBD mu=8",
            "batch": batch, "mfma_us": ms_m * 1e3, "valu_us": ms_v * 1e3, "outputs_equal": same,
            "mfma_products_per_s": batch / (ms_m / 1e3), "valu_products_per_s": batch / (ms_v / 1e3),
            "mfma_over_valu": ms_v / ms_m,
            "roofline": {"bound": "hbm", "achieved": algo / (ms_m / 1e3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": algo / (ms_m / 1e3) / 1e9 / HBM_PEAK_GBS, "algorithmic_bytes": algo,
                         "floor_us": algo / (HBM_PEAK_GBS * 1e9) * 1e6}}


def batch1_bench(ctx, stream, reps=200):
    """One Saber L=3 MatrixVectorMul with rounding (config 1's shape) per call:
    the GPU's per-call latency for a single instance (event pairs around each
    launch, inputs resident), MFMA and VALU."""
    A, s = synthetic_inputs(1, seed=14)
    dA, ds_, do = T.DeviceBuffer.from_numpy(A), T.DeviceBuffer.from_numpy(s), T.DeviceBuffer(L * N * 2)
    out = {"workload": "one Saber L=3 MatrixVectorMul + round q->p per call"}
    for name, backend in (("mfma", T.MFMA), ("valu", T.VALU)):
        run = lambda: ctx.matvec(do, dA, ds_, 1, T.SABER, backend, round_to_p=True, stream=stream)
        for _ in range(20):
            run()
        stream.synchronize()
        e0, e1 = T.Event(), T.Event()
        ts = []
        for _ in range(reps):
            e0.record(stream)
            run()
            e1.record(stream)
            e1.synchronize()
            ts.append(e0.elapsed_ms(e1) * 1e3)
        e0.destroy()
        e1.destroy()
        out[f"{name}_us_p50"] = percentile(ts, 50)
        out[f"{name}_us_p99"] = percentile(ts, 99)
    ctx.check(stream)
    for b in (dA, ds_, do):
        b.free()
    return out


def calib_copy(mib):
    """A plain copy of known size: read mib MiB, write mib MiB (runtime copy kernel).
    Its FETCH_SIZE / WRITE_SIZE readings give the scale of those counters."""
    nbytes = mib << 20
    src, dst = T.DeviceBuffer(nbytes).zero(), T.DeviceBuffer(nbytes)
    T.device_synchronize()
    for _ in range(3):
        T._hipcheck(T.hip().hipMemcpy(dst.ptr, src.ptr, nbytes, T.D2D), "hipMemcpy")
    T.device_synchronize()
    print(json.dumps({"calib_copy_bytes": nbytes, "copies": 3}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    # The device takes about 150 back-to-back launches of the B=65536 kernel
    # to settle its clocks after start-up (per-launch time 123 -> 166 -> 120 us):
    # the settle phase below runs past that transient and is reported.
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--settle", type=int, default=400,
                    help="back-to-back launches between the cold and the settled timing (reported)")
    ap.add_argument("--batch", type=int, default=BATCH)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-kem", action="store_true", help="skip the batched KEM leg")
    ap.add_argument("--kem-batch", type=int, default=16384)
    # 100 timed calls, so that p99 is the 99th of 100 and not simply the maximum
    ap.add_argument("--kem-reps", type=int, default=100)
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-parameter leg")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving leg")
    ap.add_argument("--serve-qps", default="2000,6000,12000,24000,48000",
                    help="offered request rates, comma separated (a sweep up to past saturation)")
    ap.add_argument("--serve-seconds", type=float, default=3.0,
                    help="seconds of arrivals per rate, capped at 36000 requests per rate")
    ap.add_argument("--only-kem", action="store_true",
                    help="only the batched KEM leg (no CPU rate): for profiler runs")
    ap.add_argument("--only-mixed", action="store_true",
                    help="only the mixed-parameter leg: for profiler runs")
    ap.add_argument("--only-matvec", action="store_true",
                    help="only the headline MFMA MatVec timing (cold, settle, settled): for profiler runs")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="only run device-to-device copies of this many MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path "
              "(instances are independent, so a batch would split across GPUs with no exchange step).",
              file=sys.stderr)
        sys.exit(2)

    if args.calib_copy_mib:
        calib_copy(args.calib_copy_mib)
        return
    if args.only_kem:
        ctx = T.Context(0)
        stream = T.Stream()
        print(json.dumps({"kem": kem_bench(ctx, stream, args.kem_batch, args.kem_reps, with_cpu=False)}))
        stream.destroy()
        ctx.close()
        return
    if args.only_mixed:
        ctx = T.Context(0)
        stream = T.Stream()
        print(json.dumps({"mixed": mixed_bench(ctx, stream), "graph_chain": graph_chain_bench(ctx, stream)}))
        stream.destroy()
        ctx.close()
        return
    print(f"[bench] Saber MatVec B={args.batch} steps={args.steps} warmup={args.warmup}", file=sys.stderr, flush=True)
    ctx = T.Context(0)
    stream = T.Stream()
    A, s = synthetic_inputs(args.batch)
    dA, ds_ = T.DeviceBuffer.from_numpy(A), T.DeviceBuffer.from_numpy(s)
    out_mfma = T.DeviceBuffer(args.batch * L * N * 2)
    out_valu = T.DeviceBuffer(args.batch * L * N * 2)

    def run_mfma():
        ctx.matvec(out_mfma, dA, ds_, args.batch, T.SABER, T.MFMA, transpose=False, round_to_p=True, stream=stream)

    def run_valu():
        ctx.matvec(out_valu, dA, ds_, args.batch, T.SABER, T.VALU, transpose=False, round_to_p=True, stream=stream)

    run_mfma()
    run_valu()
    stream.synchronize()
    ctx.check(stream)
    same = bool(np.array_equal(out_mfma.download(np.uint16), out_valu.download(np.uint16)))
    if not same:
        print("[bench] ERROR: MFMA and VALU outputs differ", file=sys.stderr)
        sys.exit(1)

    # Timed twice: right after the W warm-up steps (the device still in its
    # start-up clock transient) and again after a reported settle phase of
    # back-to-back launches; `value` is the settled rate, the cold one is
    # reported beside it.
    ms_cold = time_launches(run_mfma, stream, args.steps, args.warmup)
    print(f"[bench] mfma cold {ms_cold / args.steps * 1e3:.1f} us/launch", file=sys.stderr, flush=True)
    settle_ms = time_launches(run_mfma, stream, args.settle, 0) if args.settle else 0.0
    ms_mfma = time_launches(run_mfma, stream, args.steps, 0)
    print(f"[bench] mfma {ms_mfma / args.steps * 1e3:.1f} us/launch after {args.settle} settle launches",
          file=sys.stderr, flush=True)
    if args.only_matvec:
        per_launch_s = ms_mfma / args.steps / 1e3
        print(json.dumps({"metric": "Saber MatVec TMVP ops/s at B=65536 (MFMA only)", "value": args.batch / per_launch_s,
                          "unit": "MatVec/s", "ms_per_step": ms_mfma / args.steps,
                          "ms_per_step_cold": ms_cold / args.steps, "steps": args.steps, "warmup": args.warmup,
                          "settle_launches": args.settle}))
        return
    # the same launch captured in a graph: median of 25 replays, each between
    # its own pair of events (so each pays its own launch and ramp-up)
    g = ctx.capture(stream, run_mfma)
    replays = []
    for _ in range(25):
        e0, e1 = T.Event(), T.Event()
        e0.record(stream)
        g.launch(stream)
        e1.record(stream)
        e1.synchronize()
        replays.append(e0.elapsed_ms(e1) * 1e3)
        e0.destroy()
        e1.destroy()
    g.destroy()
    graph_us = sorted(replays)[len(replays) // 2]
    print(f"[bench] mfma graph replay median {graph_us:.1f} us", file=sys.stderr, flush=True)
    valu_steps = max(5, args.steps // 10)
    ms_valu = time_launches(run_valu, stream, valu_steps, max(2, args.warmup // 10))
    print(f"[bench] valu {ms_valu / valu_steps * 1e3:.1f} us/launch", file=sys.stderr, flush=True)
    ctx.check(stream)

    per_launch_s = ms_mfma / args.steps / 1e3
    value = args.batch / per_launch_s
    valu_value = args.batch / (ms_valu / valu_steps / 1e3)
    achieved = BYTES_PER_INSTANCE * args.batch / per_launch_s / 1e9
    result = {
        "metric": "Saber MatVec TMVP ops/s at B=65536 (batched MFMA vs plain VALU); KEM p99 ms",
        "value": value,
        "unit": "MatVec/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "settle_launches": args.settle,
        "settle_ms": settle_ms,
        "ms_per_step": ms_mfma / args.steps,
        "ms_per_step_cold": ms_cold / args.steps,
        "value_cold": args.batch / (ms_cold / args.steps / 1e3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8 MFMA limbs, int32 accumulate (u16 in/out)",
        "data": "synthetic (seeded uniform 13-bit A, CBD mu=8 secrets)",
        "config": {"workload": "Saber L=3 MatrixVectorMul A.s + round q->p", "batch": args.batch,
                   "param_set": "Saber (q=2^13, p=2^10)", "backend": "mfma"},
        "valu": {"value": valu_value, "unit": "MatVec/s", "ms_per_step": ms_valu / valu_steps,
                 # the plain form's multiply-adds are packed 16-bit (two per lane per instruction)
                 "mac_per_s": MACS_PER_INSTANCE * valu_value,
                 "valu_util_of_peak": MACS_PER_INSTANCE * valu_value / VALU_PEAK_MACS,
                 "valu_peak": "256 CU x 4 SIMD x 16 lanes x 2.4 GHz x 2 MACs per packed 16-bit multiply-add "
                              "= 78.6 T MAC/s"},
        "mfma_graph_replay_median_us": graph_us,
        "mfma_over_valu": value / valu_value,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic() if args.batch == BATCH else None},
        "mfma_tops_useful": MACS_PER_INSTANCE * 2 * value / 1e12,
        "outputs_equal": same,
    }
    executed = MACS_PER_INSTANCE * 2 * 2 * value / 1e12  # two int8 limbs per product
    result["mfma_executed_tops"] = executed
    result["mfma_util_of_published"] = executed / MFMA_I8_PUBLISHED_TOPS
    peaks = measured_peaks()
    if peaks:
        result["roofline"]["measured"] = {
            **peaks, "frac_of_stream": achieved / peaks["hbm_stream_GBs"],
            "mfma_executed_tops": executed, "mfma_util_of_measured": executed / peaks["mfma_i8_tops"],
            "mfma_util_of_published": executed / MFMA_I8_PUBLISHED_TOPS}
    print("[bench] raw polymul B=65536 ...", file=sys.stderr, flush=True)
    result["polymul"] = polymul_bench(ctx, stream, args.steps)
    print(f"[bench] polymul mfma {result['polymul']['mfma_us']:.1f} us, valu {result['polymul']['valu_us']:.1f} us",
          file=sys.stderr, flush=True)
    result["batch1"] = batch1_bench(ctx, stream)
    if not args.no_mixed:
        print("[bench] mixed-parameter batch ...", file=sys.stderr, flush=True)
        result["mixed"] = mixed_bench(ctx, stream)
        print(f"[bench] mixed: separate {result['mixed']['separate_launches_ms'] * 1e3:.1f} us, persistent "
              f"{result['mixed']['persistent_autotuned_ms'] * 1e3:.1f} us, graph "
              f"{result['mixed']['graph_replay_ms'] * 1e3:.1f} us", file=sys.stderr, flush=True)
        result["graph_chain"] = graph_chain_bench(ctx, stream)
        print(f"[bench] 32 batch-1 launches: eager {result['graph_chain']['eager_ms'] * 1e3:.1f} us, graph "
              f"{result['graph_chain']['graph_replay_ms'] * 1e3:.1f} us", file=sys.stderr, flush=True)
    if not args.no_kem:
        print(f"[bench] KEM B={args.kem_batch} reps={args.kem_reps} ...", file=sys.stderr, flush=True)
        result["kem"] = kem_bench(ctx, stream, args.kem_batch, args.kem_reps, with_cpu=not args.no_cpu)
        kp = keccak_peak()
        if kp:
            result["kem"]["roofline"] = {"bound": "valu (Keccak-f[1600])", "achieved": result["kem"]["keccak_f_per_s"],
                                         "peak": kp["perms_per_s"], "unit": "Keccak-f/s",
                                         "frac": result["kem"]["keccak_f_per_s"] / kp["perms_per_s"],
                                         "peak_shape": kp["shape"], "peak_by_shape": kp["by_shape"]}
        print(f"[bench] KEM encaps+decaps p99 {result['kem']['encaps_decaps_ms']['p99']:.2f} ms", file=sys.stderr,
              flush=True)
        # every parameter set at the same batch (20 timed calls each, round trip checked)
        result["kem_all_sets"] = {}
        for ps in ("lightsaber", "saber", "firesaber", "sable1", "sable2", "sable3"):
            r = kem_bench(ctx, stream, args.kem_batch, 20, ps=ps, with_cpu=False)
            result["kem_all_sets"][ps] = {
                "keypair_ms_p50": r["keypair_ms"]["p50"], "encaps_ms_p50": r["encaps_ms"]["p50"],
                "decaps_ms_p50": r["decaps_ms"]["p50"], "keccak_f_per_s": r["keccak_f_per_s"],
                "encaps_decaps_per_s": r["encaps_decaps_per_s"]}
        print("[bench] KEM all sets: " + ", ".join(f"{k} {v['encaps_ms_p50'] * 1e3:.0f} us" for k, v in
                                                   result["kem_all_sets"].items()), file=sys.stderr, flush=True)
    if not args.no_serving:
        result["serving"] = []
        for qps in (float(x) for x in str(args.serve_qps).split(",")):
            print(f"[bench] serving {qps:.0f} req/s ...", file=sys.stderr, flush=True)
            sv = serving_bench(ctx, qps, min(args.serve_seconds, 36000.0 / qps))
            result["serving"].append(sv)
            print(f"[bench] serving p50 {sv['p50_ms']:.2f} ms p99 {sv['p99_ms']:.2f} ms, "
                  f"{sv['instances_per_s'] / 1e6:.2f} M instances/s", file=sys.stderr, flush=True)
    if not args.no_cpu:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        result["cpu_baseline"] = cpu_baseline()
    stream.destroy()
    ctx.close()
    print(json.dumps(result))


if __name__ == "__main__":
    main()
