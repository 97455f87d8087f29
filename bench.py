"""Benchmark of the plain commit pipeline on one MI355X.

One step = one full proof of the plain GPU workload: a single Fibonacci M31
column of 2^24 rows already resident in HBM -> circle LDE on the 2^26 domain
-> Blake2s trace commitment -> quotient commitment -> circle fold and 22 line
folds, each fused with its Blake2s commitment -> Fiat-Shamir queries (100)
-> decommitment, leaving the KRAK v1 proof bytes in HBM. All steps are
enqueued back to back on the prover's stream and timed between device
synchronisations.

Prints one JSON line (see README of the contract in DESIGN.md).

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--log-n 24] [--no-cpu]
"""

from __future__ import annotations

import argparse
import ctypes as C
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

METRIC = "LDE+commit+FRI Melem/s (frac of HBM roofline); batched proofs/s, p50/p99 lat"
HBM_PEAK_GBS = 8000.0
# int32 VALU issue rate: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz (lane-ops/s)
VALU_PEAK_TOPS = 256 * 4 * 16 * 2.4e9 / 1e12
# VALU ops of one Blake2s compression: 10 rounds x 8 G x 12 ops + 16 feed-forward xors
OPS_PER_COMPRESSION = 10 * 8 * 12 + 16


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def golden_case(label):
    path = os.path.join(ROOT, "tests", "golden", "proofs.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for c in json.load(f):
            if c["label"] == label:
                return c
    return None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--log-n", type=int, default=24)
    ap.add_argument("--queries", type=int, default=100)
    ap.add_argument("--log-last", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-log-n", type=int, default=24)
    ap.add_argument("--cpu-reps", type=int, default=2)
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="also run a device-to-device copy of this many MiB (counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(independent proofs shard by job; see DESIGN.md). Run with --gpus 1.", file=sys.stderr)
        return 2

    import vortexstark_amd as vx

    assert vx.gpu_available(), "no HIP device"
    log_n, nq, last = args.log_n, args.queries, args.log_last
    golden = golden_case("gpu_plain") if (log_n, nq, last) == (24, 100, 4) else None
    if golden:
        a, b = golden["a"], golden["b"]
    else:
        a, b = 1, 1
    log(f"bench: log_n={log_n} queries={nq} last=2^{last} a={a} b={b}")
    t0 = time.time()
    prover = vx.Prover(log_n, nq, last)
    log(f"prover ready ({time.time() - t0:.2f}s: twiddles + workspace)")
    trace = vx.fibonacci_trace_device(a, b, log_n)
    stream = C.c_void_p(prover.stream)

    for _ in range(args.warmup):
        prover.prove_device(trace.addr, a, b)
    vx.lib.vx_stream_synchronize(stream)
    vx.check("warmup")
    log(f"warmup done ({args.warmup} proofs)")
    if args.calibrate_copy_mb:
        nb = args.calibrate_copy_mb << 20
        src, dst = vx.DeviceBuffer(nb // 4), vx.DeviceBuffer(nb // 4)
        vx.lib.vx_memcpy_d2d(dst.ptr, src.ptr, nb)
        vx.synchronize()
        log(f"calibration copy: {args.calibrate_copy_mb} MiB device to device")
        src.free()
        dst.free()

    vx.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        prover.prove_device(trace.addr, a, b)
    vx.lib.vx_stream_synchronize(stream)
    vx.synchronize()
    elapsed = time.perf_counter() - t_start
    vx.check("timed steps")
    ms_per_step = 1e3 * elapsed / args.steps
    n_rows = 1 << log_n
    melem_s = n_rows * args.steps / elapsed / 1e6
    log(f"timed: {args.steps} proofs in {elapsed:.3f}s -> {ms_per_step:.3f} ms/proof, {melem_s:.1f} Melem/s")

    proof = prover.proof_bytes()
    parity = None
    if golden:
        parity = hashlib.sha256(proof).hexdigest() == golden["sha256"]
        log(f"proof bytes identical to the oracle's golden summary: {parity}")

    # single-proof latency (each proof synchronised on its own)
    lat = []
    for _ in range(max(5, min(args.steps, 20))):
        t = time.perf_counter()
        prover.prove_device(trace.addr, a, b)
        vx.lib.vx_stream_synchronize(stream)
        lat.append(1e3 * (time.perf_counter() - t))
    lat.sort()
    p50 = lat[len(lat) // 2]
    p99 = lat[min(len(lat) - 1, int(round(0.99 * (len(lat) - 1))))]

    # launch timeline of one proof (HIP events on the prover's stream)
    vx.lib.vx_prover_set_timing(prover.handle, 1)
    prover.prove_device(trace.addr, a, b)
    vx.lib.vx_stream_synchronize(stream)
    phases = {}
    for i in range(1, vx.lib.vx_prover_timeline_count(prover.handle)):
        label = vx.lib.vx_prover_timeline_label(prover.handle, i).decode()
        ms = vx.lib.vx_prover_timeline_ms(prover.handle, i)
        phases[label] = phases.get(label, 0.0) + ms
    vx.lib.vx_prover_set_timing(prover.handle, 0)
    vx.check("timeline")

    E = 1 << (log_n + 2)
    # dominant kernel: the first commit stage of the quotient tree (Blake2s over E leaves)
    dom = "quotient_commit:stage_leaf"
    dom_ms = phases.get(dom)
    roofline = None
    if dom_ms:
        logc = 5
        while logc > 0 and (E >> (7 + logc)) < 1024:
            logc -= 1
        compressions = E + (E - (E >> (logc + 1)))  # leaf hashes + nodes inside the stage
        ops = compressions * OPS_PER_COMPRESSION
        achieved = ops / (dom_ms * 1e-3) / 1e12
        roofline = {"bound": "valu", "kernel": "merkle_stage_kernel<SrcQuotient>", "achieved": round(achieved, 3),
                    "peak": round(VALU_PEAK_TOPS, 2), "unit": "Tops/s (int32 lane-ops)",
                    "frac": round(achieved / VALU_PEAK_TOPS, 4), "traffic": None,
                    "avg_ms": round(dom_ms, 4), "ops_per_launch": ops}
    ntt_ms = phases.get("lde:evaluate")
    roofline_hbm = None
    if ntt_ms:
        n = 1 << log_n
        byts = 4 * n + 4 * E  # read coefficients once, write the LDE once
        achieved = byts / (ntt_ms * 1e-3) / 1e9
        roofline_hbm = {"bound": "hbm", "kernel": "ntt_pass_kernel<true> x3 (evaluate 2^%d)" % (log_n + 2),
                        "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "phase_ms": round(ntt_ms, 4)}
    # pipeline-level bound: 6E Blake2s compressions at the VALU peak
    total_compressions = 6 * E
    t_valu_ms = total_compressions * OPS_PER_COMPRESSION / (VALU_PEAK_TOPS * 1e12) * 1e3

    cpu = None
    if not args.no_cpu:
        import oracle_lib

        threads = min(16, os.cpu_count() or 1)
        L = oracle_lib.lib()
        L.orc_set_threads(threads)
        cl = args.cpu_log_n
        log(f"cpu baseline: oracle prove at log_n={cl}, {args.cpu_reps} reps, {threads} threads")
        secs = L.orc_time_prove_fibonacci(a % vx.P, b % vx.P, cl, nq, last, args.cpu_reps)
        cpu_rate = (1 << cl) * args.cpu_reps / secs / 1e6
        model = ""
        try:
            with open("/proc/cpuinfo") as f:
                model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
        except OSError:
            pass
        cpu = {"value": round(cpu_rate, 4), "unit": "Melem/s", "cores": threads, "kind": "port",
               "sample": f"oracle (C++ restatement) full proof of a 2^{cl}-row Fibonacci column, {nq} queries, "
                         f"last layer 2^{last}, {args.cpu_reps} reps after twiddle precompute, {secs:.2f}s; {model}"}
        log(f"cpu baseline: {cpu_rate:.3f} Melem/s ({secs:.2f}s)")

    out = {
        "metric": METRIC,
        "value": round(melem_s, 2),
        "unit": "Melem/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32 (exact M31/QM31)",
        "data": "synthetic",
        "config": {"workload": "plain: Fibonacci M31 column 2^%d rows, blowup 4 (2^%d LDE), Blake2s Merkle, "
                               "FRI to 2^%d, %d queries, full proof in HBM" % (log_n, log_n + 2, last, nq),
                   "log_n": log_n, "blowup": 4, "n_queries": nq, "log_last_layer": last, "parallelism": "none"},
        "proofs_per_s": round(args.steps / elapsed, 3),
        "latency_ms": {"p50": round(p50, 3), "p99": round(p99, 3)},
        "roofline": roofline,
        "roofline_hbm": roofline_hbm,
        "pipeline_valu_bound_ms": round(t_valu_ms, 3),
        "pipeline_frac_of_valu_bound": round(t_valu_ms / ms_per_step, 4),
        "phases_ms": {k: round(v, 4) for k, v in phases.items()},
        "parity_vs_oracle": parity,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    prover.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
