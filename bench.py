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

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
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
        "metric": "Saber MatVec TMVP ops/s at B=65536 (batched MFMA vs plain VALU)",
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
    if not args.no_cpu:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        result["cpu_baseline"] = cpu_baseline()
    stream.destroy()
    ctx.close()
    print(json.dumps(result))


if __name__ == "__main__":
    main()
