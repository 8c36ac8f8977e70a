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
`cpu_baseline` times the CPU oracle (a C restatement of the reference's
Toom-4/schoolbook path) on a bounded sample of the same workload.

`kem` reports the full batched Saber KEM at 16384 instances (key generation,
encapsulation, decapsulation with all sampling, hashing, codecs and products on
the GPU; seeds and keys resident in HBM): ops/s per step and p50/p99 of the
GPU time of each batched call, after a round-trip check of every instance.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-kem] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "asfand-tmvp_amd"))

import tmvp_amd as T  # noqa: E402

N = 256
L = 3
BATCH = 65536
# Algorithmic HBM bytes per instance: A (9 x 256 x 2) + s (3 x 256 x 2) + out (3 x 256 x 2).
BYTES_PER_INSTANCE = 9 * N * 2 + 3 * N * 2 + 3 * N * 2
MACS_PER_INSTANCE = 9 * N * N
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


def cpu_baseline(seconds=12.0):
    """Time the CPU oracle (Toom-4 over Karatsuba, the faster of the two CPU
    restatements) on host cores over a bounded sample of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # test infrastructure: used only as the timed CPU baseline

    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    A, s = synthetic_inputs(4096, seed=2)
    O.matvec_batch(L, A[:64], s[:64], method=1, nthreads=cores)  # warm-up
    done, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        O.matvec_batch(L, A, s, method=1, nthreads=cores)
        done += A.shape[0]
    dt = time.perf_counter() - t0
    return {"value": done / dt, "unit": "MatVec/s", "cores": cores, "kind": "port",
            "sample": f"{done} Saber L=3 MatrixVectorMul instances (Toom-4, {cores} threads, {dt:.1f} s)"}


def percentile(xs, q):
    xs = sorted(xs)
    k = min(len(xs) - 1, max(0, int(np.ceil(q / 100.0 * len(xs))) - 1))
    return xs[k]


def kem_bench(ctx, stream, batch=16384, reps=40, ps="saber", cpu_seconds=4.0, with_cpu=True):
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
    for _ in range(3):
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
    out = {
        "workload": f"Saber KEM keypair/encaps/decaps, {batch} instances per call",
        "batch": batch, "reps": reps,
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

        cores = max(1, min(len(os.sched_getaffinity(0)), 16))
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
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--batch", type=int, default=BATCH)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-kem", action="store_true", help="skip the batched KEM leg")
    ap.add_argument("--kem-batch", type=int, default=16384)
    ap.add_argument("--kem-reps", type=int, default=40)
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="only run device-to-device copies of this many MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(instances are independent; see DESIGN.md for how B would shard).", file=sys.stderr)
        sys.exit(2)

    if args.calib_copy_mib:
        calib_copy(args.calib_copy_mib)
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

    ms_mfma = time_launches(run_mfma, stream, args.steps, args.warmup)
    print(f"[bench] mfma {ms_mfma / args.steps * 1e3:.1f} us/launch", file=sys.stderr, flush=True)
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
        "ms_per_step": ms_mfma / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8 MFMA limbs, int32 accumulate (u16 in/out)",
        "data": "synthetic (seeded uniform 13-bit A, CBD mu=8 secrets)",
        "config": {"workload": "Saber L=3 MatrixVectorMul A.s + round q->p", "batch": args.batch,
                   "param_set": "Saber (q=2^13, p=2^10)", "backend": "mfma"},
        "valu": {"value": valu_value, "unit": "MatVec/s", "ms_per_step": ms_valu / valu_steps},
        "mfma_over_valu": value / valu_value,
        "polymul_per_s": value * 9,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic() if args.batch == BATCH else None},
        "mfma_tops_useful": MACS_PER_INSTANCE * 2 * value / 1e12,
        "outputs_equal": same,
    }
    if not args.no_kem:
        print(f"[bench] KEM B={args.kem_batch} reps={args.kem_reps} ...", file=sys.stderr, flush=True)
        result["kem"] = kem_bench(ctx, stream, args.kem_batch, args.kem_reps, with_cpu=not args.no_cpu)
        print(f"[bench] KEM encaps+decaps p99 {result['kem']['encaps_decaps_ms']['p99']:.2f} ms", file=sys.stderr,
              flush=True)
    if not args.no_cpu:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        result["cpu_baseline"] = cpu_baseline()
    stream.destroy()
    ctx.close()
    print(json.dumps(result))


if __name__ == "__main__":
    main()
