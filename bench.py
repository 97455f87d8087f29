"""Benchmark of the commit pipeline on one MI355X.

Headline (value): the plain GPU workload. One step = one full proof of a
single Fibonacci M31 column of 2^24 rows already resident in HBM -> circle LDE
on the 2^26 domain -> Blake2s trace commitment -> quotient commitment ->
circle fold and 22 line folds, each fused with its Blake2s commitment ->
Fiat-Shamir queries (100) -> decommitment, leaving the KRAK v1 proof bytes in
HBM. All steps are enqueued back to back on the prover's stream and timed
between device synchronisations.

The same JSON line carries the other workloads of the metric:
  wide    34 uniform M31 columns (4 x 2^14, 10 x 2^18, 20 x 2^20), stwo LDE
          blowup 4, one lifted Blake2s tree, fused FRI fold+commit to 2^4;
  batch   256 jobs of 1-8 Fibonacci columns, log_n uniform in [10, 18]:
          batched proving (graph replay) vs the same proofs one at a time;
  serve   Poisson arrivals of jobs with 1-34 columns, log_n uniform in
          [12, 26], admitted under the HBM budget: proofs/s and p50/p99;
  host    the plain workload with its trace in host memory (pinned and
          pageable), uploads overlapping the previous proof;
  deep    stwo DEEP quotient of the wide shape's 34 LDE columns at two
          sample points on the 2^22 lifting domain;
  grind   256 Blake2s proof-of-work searches at 26 bits from channel
          states on the device: nonces/s against the VALU roofline.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--no-cpu]
                    [--skip wide,batch,serve,host,deep,grind]
"""

from __future__ import annotations

import argparse
import ctypes as C
import hashlib
import json
import os
import sys
import time

import numpy as np

# Hardware queues the HIP runtime gives this process (read once, when it starts).
# The batched and served workloads keep up to two provers per size class, each
# on its own stream: with HIP's default of 4 queues, streams share queues and a
# short batch waits behind a long one dispatched ahead of it on the same queue.
# A smaller setting in the environment is raised to 16.
if not os.environ.get("GPU_MAX_HW_QUEUES", "").isdigit() or int(os.environ["GPU_MAX_HW_QUEUES"]) < 16:
    os.environ["GPU_MAX_HW_QUEUES"] = "16"

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

METRIC = "LDE+commit+FRI Melem/s (frac of HBM roofline); batched proofs/s, p50/p99 lat"
HBM_PEAK_GBS = 8000.0
# int32 VALU issue rate: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz (lane-ops/s)
VALU_PEAK_TOPS = 256 * 4 * 16 * 2.4e9 / 1e12
# VALU ops of one Blake2s compression on gfx950: 10 rounds x 8 G x 12 ops (v_add3,
# v_add, v_xor, v_alignbit) + 8 three-input xors of the feed-forward (v_bitop3)
OPS_PER_COMPRESSION = 10 * 8 * 12 + 8
# one radix-2 M31 butterfly in the pass kernel's gfx950 ISA: product
# (v_mad_u64_u32, alignbit, and, add, add of -p, min), modular add (add, add of
# -p, min) and modular subtract (sub, add of p, min)
OPS_PER_BUTTERFLY = 12
# VALU ops of one nonce of the proof-of-work search for pow_bits <= 32 (the loop
# body of grind_kernel in its gfx950 ISA: 279 v_xor, 279 v_alignbit, 280 adds
# and the check): the uniform part of round 0 runs once per wave and only digest
# word 0 is formed
OPS_PER_NONCE = 847
POW_BITS = 26  # the reference transcripts' PoW (prover.rs:96)
# plain proofs in flight for the pipelined figure (2^24 rows: 2 -> 11.27 ms,
# 3 -> 11.09 ms, 4 -> 11.29 ms per proof, measured on MI355X)
PIPE_PROVERS = 3
# VALU lane-ops the DEEP kernels issue, counted with rocprofv3 (SQ_INSTS_VALU x 64
# per dispatch, profiles/deep_quotient_wide_shape_and_pow_grind_26bit_kernel_stats.md):
# quotients_combine_kernel<8> per (row, accumulation); deep_fused_kernel<2,4> per
# row of the wide shape's lifting domain (34 columns, 2 points). The fused figure
# is the count of its form without the next-group prefetch (the same arithmetic
# in fewer instructions); the prefetching form issues 1491 and runs faster.
OPS_PER_ROW_ACC = 259
FUSED_OPS_PER_ROW = 1388


def load_traffic() -> dict:
    """HBM bytes per launch of the rooflined kernels, measured offline with
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs (scaled by a copy-kernel calibration);
    see the summary the file names."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_per_launch.json")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f).get("kernels", {})


TRAFFIC = load_traffic()


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


def stage_compressions(n_items: int, leaf_blocks: int) -> int:
    """Blake2s compressions of the first stage kernel of a tree: leaf hashes of
    n_items leaves (leaf_blocks each) and every node below each wave's output
    (a wave reduces 128 * 2^logc items to one node; logc as the library picks it)."""
    logc = 5
    while logc > 0 and (n_items >> (7 + logc)) < 1024:
        logc -= 1
    waves = n_items >> (7 + logc)
    return n_items * leaf_blocks + (n_items - waves)


MEASURED = {}  # filled at start-up: measured VALU peak, NTT plans


def valu_roofline(kernel: str, compressions: int, ms: float) -> dict:
    ops = compressions * OPS_PER_COMPRESSION
    achieved = ops / (ms * 1e-3) / 1e12
    out = {"bound": "valu", "kernel": kernel, "achieved": round(achieved, 3), "peak": round(VALU_PEAK_TOPS, 2),
           "unit": "Tops/s (int32 lane-ops)", "frac": round(achieved / VALU_PEAK_TOPS, 4),
           "traffic": TRAFFIC.get(kernel), "avg_ms": round(ms, 4), "ops_per_launch": ops}
    if "valu_peak" in MEASURED:
        out["peak_measured"] = round(MEASURED["valu_peak"], 3)
    return out


def cpu_model() -> str:
    """The host CPU as /proc/cpuinfo names it: model name, family/model/stepping,
    logical CPUs listed."""
    try:
        with open("/proc/cpuinfo") as f:
            text = f.read()
    except OSError:
        return ""
    first = {}
    for ln in text.splitlines():
        if ":" in ln:
            k, v = (x.strip() for x in ln.split(":", 1))
            first.setdefault(k, v)
    n = sum(1 for ln in text.splitlines() if ln.startswith("processor"))
    return (f"{first.get('model name', '?')} (vendor {first.get('vendor_id', '?')}, family {first.get('cpu family', '?')}, "
            f"model {first.get('model', '?')}, stepping {first.get('stepping', '?')}, {n} logical CPUs listed)")


def host_threads() -> tuple:
    """Threads for the CPU baseline: every core this process may use (its
    affinity mask, capped by a cgroup CPU quota and by OMP_NUM_THREADS when
    either is set: a shared host grants each GPU a share of its cores), with how
    the count was found."""
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = os.cpu_count() or 1
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, period = f.read().split()[:2]
        if q != "max":
            quota = float(q) / float(period)
    except (OSError, ValueError):
        pass
    threads = affinity if quota is None else max(1, min(affinity, int(quota + 0.5)))
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        threads = min(threads, int(omp))
    info = {"affinity_cpus": affinity, "cgroup_cpu_quota": quota, "omp_num_threads": omp or None,
            "os_cpu_count": os.cpu_count(), "threads_used": threads, "cpu": cpu_model()}
    return threads, info


HOST_THREADS, HOST_INFO = host_threads()


def quantiles(xs):
    xs = sorted(xs)
    return {"p50": round(xs[len(xs) // 2], 3), "p99": round(xs[min(len(xs) - 1, int(round(0.99 * (len(xs) - 1))))], 3)}


def bench_plain(vx, args) -> dict:
    log_n, nq, last = args.log_n, args.queries, args.log_last
    golden = golden_case("gpu_plain") if (log_n, nq, last) == (24, 100, 4) else None
    if golden:
        a, b = golden["a"], golden["b"]
    else:
        a, b = 1, 1
    log(f"plain: log_n={log_n} queries={nq} last=2^{last} a={a} b={b}")
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
    calibration_copy(vx, args)

    vx.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        prover.prove_device(trace.addr, a, b)
    vx.lib.vx_stream_synchronize(stream)
    vx.synchronize()
    elapsed = time.perf_counter() - t_start
    vx.check("timed steps")
    ms_per_step = 1e3 * elapsed / args.steps
    melem_s = (1 << log_n) * args.steps / elapsed / 1e6
    log(f"timed: {args.steps} proofs in {elapsed:.3f}s -> {ms_per_step:.3f} ms/proof, {melem_s:.1f} Melem/s")

    proof = prover.proof_bytes()
    parity = None
    if golden:
        parity = hashlib.sha256(proof).hexdigest() == golden["sha256"]
        log(f"proof bytes identical to the oracle's golden summary: {parity}")

    # the same proofs with PIPE_PROVERS in flight (one prover and stream each):
    # one proof's latency-bound FRI chain overlaps the others' hashing
    pipelined = None
    if not args.no_pipelined:
        extra = [vx.Prover(log_n, nq, last) for _ in range(PIPE_PROVERS - 1)]
        traces = [vx.fibonacci_trace_device(a, b, log_n) for _ in extra]
        ring = [(prover, trace)] + list(zip(extra, traces))
        for pv, tr in ring:
            pv.prove_device(tr.addr, a, b)
        vx.synchronize()
        t = time.perf_counter()
        for i in range(args.steps):
            pv, tr = ring[i % PIPE_PROVERS]
            pv.prove_device(tr.addr, a, b)
        vx.synchronize()
        el2 = time.perf_counter() - t
        vx.check("pipelined steps")
        same = all(pv.proof_bytes() == proof for pv in extra)
        pipelined = {"streams": PIPE_PROVERS, "ms_per_proof": round(1e3 * el2 / args.steps, 4),
                     "melem_s": round((1 << log_n) * args.steps / el2 / 1e6, 2), "identical_proof": same}
        log(f"pipelined ({PIPE_PROVERS} streams): {1e3 * el2 / args.steps:.3f} ms/proof, "
            f"{(1 << log_n) * args.steps / el2 / 1e6:.1f} Melem/s, identical proof: {same}")
        for pv in extra:
            pv.close()
        for tr in traces:
            tr.free()

    # single-proof latency (each proof synchronised on its own)
    lat = []
    for _ in range(max(5, min(args.steps, 20))):
        t = time.perf_counter()
        prover.prove_device(trace.addr, a, b)
        vx.lib.vx_stream_synchronize(stream)
        lat.append(1e3 * (time.perf_counter() - t))

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
    dom_ms = phases.get("quotient_commit:stage_leaf")
    roofline = valu_roofline("merkle_stage_kernel<SrcQuotient>", stage_compressions(E, 1), dom_ms) if dom_ms else None
    ntt_ms = phases.get("lde:evaluate")
    # the LDE's evaluation plan; when it ends in a replicating pass the library
    # runs the interpolation's layers above the pass's start in that pass too
    # (the LDE bridge), so the evaluate phase also holds those layers
    lde_plan = vx.ntt_lde_plan(log_n + 2)
    bridged = bool(lde_plan) and lde_plan[-1][2] == 1
    bridged_layers = log_n - sum(m for m, _, _ in lde_plan[:-1]) if bridged else 0
    what = "evaluate 2^%d, all passes" % (log_n + 2)
    if bridged:
        what += " + the interpolation's top %d layers (LDE bridge)" % bridged_layers
    # counted traffic belongs to the plan it was measured with
    lde_traffic = next((t for t in TRAFFIC.get("lde:evaluate", [])
                        if [list(x) for x in t.get("plan", [])] == [list(x) for x in lde_plan]), None)
    roofline_hbm = None
    if ntt_ms:
        byts = 4 * (1 << log_n) + 4 * E  # read the quarter once, write the LDE once
        achieved = byts / (ntt_ms * 1e-3) / 1e9
        roofline_hbm = {"bound": "hbm", "kernel": "ntt_pass_kernel<true> (%s)" % what,
                        "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": lde_traffic,
                        "phase_ms": round(ntt_ms, 4), "passes": lde_plan}
    # the same phase priced against the VALU: its butterflies are the work
    # (the top two layers of a zero-padded input only copy and are not counted)
    roofline_ntt = None
    if ntt_ms:
        k = log_n + 2
        bfly = (k - 2) * (E // 2) + bridged_layers * ((1 << log_n) // 2)
        ops = bfly * OPS_PER_BUTTERFLY
        achieved = ops / (ntt_ms * 1e-3) / 1e12
        roofline_ntt = {"bound": "valu", "kernel": "ntt_pass_kernel<true> (%s)" % what,
                        "achieved": round(achieved, 3), "peak": VALU_PEAK_TOPS, "unit": "Tops/s (int32 lane-ops)",
                        "frac": round(achieved / VALU_PEAK_TOPS, 4), "butterflies": bfly,
                        "ops_per_butterfly": OPS_PER_BUTTERFLY, "phase_ms": round(ntt_ms, 4)}
    # pipeline-level bound: 6E Blake2s compressions at the VALU peak
    t_valu_ms = 6 * E * OPS_PER_COMPRESSION / (VALU_PEAK_TOPS * 1e12) * 1e3

    cpu = None
    if not args.no_cpu:
        import oracle_lib

        threads = HOST_THREADS
        L = oracle_lib.lib()
        L.orc_set_threads(threads)
        cl = args.cpu_log_n
        log(f"cpu baseline: oracle prove at log_n={cl}, 1 warm-up + {args.cpu_reps} timed, {threads} threads")
        each = np.zeros(args.cpu_reps, np.float64)
        L.orc_time_prove_each(a % vx.P, b % vx.P, cl, nq, last, 1, args.cpu_reps, each.ctypes.data)
        secs = float(np.median(each))
        cpu_rate = (1 << cl) / secs / 1e6
        cpu = {"value": round(cpu_rate, 4), "unit": "Melem/s", "cores": threads, "kind": "port",
               "sample": f"oracle (C++ restatement) full proof of the same 2^{cl}-row Fibonacci column, {nq} queries, "
                         f"last layer 2^{last}: median of {args.cpu_reps} proofs after one warm-up "
                         f"({', '.join(f'{x:.2f}' for x in each)} s), twiddles precomputed",
               "host": HOST_INFO}
        log(f"cpu baseline: {cpu_rate:.3f} Melem/s (median {secs:.2f}s)")
        # the reference's CPU-runnable case: 2^16 rows, 64 queries, FRI to 2^4
        c1 = np.zeros(7, np.float64)
        L.orc_time_prove_each(a % vx.P, b % vx.P, 16, 64, 4, 1, len(c1), c1.ctypes.data)
        c1_ms = 1e3 * float(np.median(c1))
        gp = vx.Prover(16, 64, 4)
        gp.prove(a, b)
        g_ms = []
        for _ in range(7):
            t = time.perf_counter()
            gp.prove(a, b)
            g_ms.append(1e3 * (time.perf_counter() - t))
        gp.close()
        cpu["config_2p16"] = {"workload": "one Fibonacci column of 2^16 rows, blowup 4 (2^18), FRI to 2^4, "
                                          "64 queries drawn by Fiat-Shamir",
                              "cpu_ms_median": round(c1_ms, 3), "cpu_runs": 7, "cpu_threads": threads,
                              "cpu_melem_s": round((1 << 16) / c1_ms / 1e3, 4),
                              "gpu_ms_median": round(float(np.median(g_ms)), 4),
                              "gpu_note": "vx_prove_fibonacci: device trace generation, proof, copy of the proof "
                                          "bytes to the host, one call at a time"}
        log(f"config 2^16: cpu median {c1_ms:.2f} ms, gpu median {np.median(g_ms):.3f} ms")
        # one thread, on a smaller trace so the run stays short
        l1 = min(cl, 18)
        L.orc_set_threads(1)
        one = np.zeros(3, np.float64)
        L.orc_time_prove_each(a % vx.P, b % vx.P, l1, nq, last, 0, len(one), one.ctypes.data)
        secs1 = float(np.median(one))
        cpu["single_thread"] = {"value": round((1 << l1) / secs1 / 1e6, 4), "unit": "Melem/s", "cores": 1,
                                "sample": f"oracle proof of a 2^{l1}-row column, median of 3, {secs1:.2f}s"}
        L.orc_set_threads(threads)
        log(f"cpu baseline, 1 thread: {(1 << l1) / secs1 / 1e6:.3f} Melem/s at 2^{l1} rows ({secs1:.2f}s)")
    prover.close()
    return {"melem_s": melem_s, "ms_per_step": ms_per_step, "elapsed": elapsed, "proofs_per_s": args.steps / elapsed,
            "latency": quantiles(lat), "roofline": roofline, "roofline_hbm": roofline_hbm, "roofline_ntt": roofline_ntt,
            "pipelined": pipelined,
            "t_valu_ms": t_valu_ms, "phases": phases, "parity": parity, "cpu": cpu}


def bench_wide(vx, args) -> dict:
    import inputs

    logs = inputs.WIDE_LOG_SIZES
    cols = inputs.m31_columns(inputs.WIDE_SEED, logs)
    w = vx.WideCommit(logs, 2, 4)
    bufs = [vx.DeviceBuffer.from_host(c) for c in cols]
    addrs = [b.addr for b in bufs]
    for _ in range(3):
        w.commit_device(addrs)
    vx.synchronize_stream(w.stream)
    K = 10
    t = time.perf_counter()
    for _ in range(K):
        w.commit_device(addrs)
    vx.synchronize_stream(w.stream)
    dt = (time.perf_counter() - t) / K
    root, fri, last, beta = w.results()
    g = None
    path = os.path.join(ROOT, "tests", "golden", "wide.json")
    if os.path.exists(path):
        with open(path) as f:
            g = next((c for c in json.load(f) if c["label"] == "wide_config"), None)
    parity = None if g is None else (root.tolist() == g["root"] and fri.tolist() == g["fri_roots"]
                                     and last.tolist() == g["last_layer"])
    w.set_timing(True)
    w.commit_device(addrs)
    vx.synchronize_stream(w.stream)
    phases = w.timeline()
    w.set_timing(False)
    trace_elems = sum(1 << l for l in logs)
    S = w.lifting_log_size
    dom_ms = phases.get("lifted_commit:stage_leaf")
    blocks = (len(logs) + 15) // 16
    roofline = valu_roofline("merkle_stage_kernel<SrcLifted>", stage_compressions(1 << S, blocks), dom_ms) \
        if dom_ms else None
    log(f"wide: {1e3 * dt:.3f} ms per commitment, {trace_elems / dt / 1e6:.1f} Melem/s, parity {parity}")
    cpu = None
    if not args.no_cpu:
        import oracle_lib

        threads = HOST_THREADS
        L = oracle_lib.lib()
        L.orc_set_threads(threads)
        ptrs = (C.c_void_p * len(cols))(*[c.ctypes.data for c in cols])
        lg = np.array(logs, np.uint32)
        secs = L.orc_time_wide_commit(ptrs, lg.ctypes.data, len(cols), 2, 4)
        cpu = {"value": round(trace_elems / secs / 1e6, 4), "unit": "Melem/s", "cores": threads, "kind": "port",
               "sample": f"oracle wide commitment of the same 34 columns, one run, {secs:.2f}s"}
        log(f"wide cpu baseline: {trace_elems / secs / 1e6:.2f} Melem/s ({secs:.2f}s)")
    w.close()
    return {"workload": "34 uniform M31 columns (4x2^14, 10x2^18, 20x2^20, splitmix64 seed 3), stwo LDE blowup 4, "
                        "lifted Blake2s tree over all columns, FRI on the beta-combination to 2^4",
            "ms_per_commit": round(1e3 * dt, 4), "melem_s": round(trace_elems / dt / 1e6, 2),
            "lde_melem_s": round(sum(1 << (l + 2) for l in logs) / dt / 1e6, 2), "parity_vs_oracle": parity,
            "roofline": roofline, "phases_ms": {k: round(v, 4) for k, v in phases.items()}, "cpu_baseline": cpu}


def batch_jobs(seed: int, n_jobs: int, min_log: int, max_log: int, max_cols: int):
    import inputs

    r = inputs.splitmix64(seed, 2 * n_jobs)
    jobs, pos = [], 0
    for j in range(n_jobs):
        log_n = min_log + int(r[2 * j] % np.uint64(max_log - min_log + 1))
        cols = 1 + int(r[2 * j + 1] % np.uint64(max_cols))
        v = inputs.splitmix64(seed + 1000, 2 * cols, pos) % np.uint64(vx_p())
        pos += 2 * cols
        jobs.append((log_n, [(int(v[2 * c]), int(v[2 * c + 1])) for c in range(cols)]))
    return jobs


def calibration_copy(vx, args) -> None:
    """A device-to-device copy of known size, the scale for the HBM byte counters."""
    if args.calibrate_copy_mb:
        nb = args.calibrate_copy_mb << 20
        src, dst = vx.DeviceBuffer(nb // 4), vx.DeviceBuffer(nb // 4)
        vx.lib.vx_memcpy_d2d(dst.ptr, src.ptr, nb)
        vx.synchronize()
        log(f"calibration copy: {args.calibrate_copy_mb} MiB device to device")
        src.free()
        dst.free()


def vx_p() -> int:
    return (1 << 31) - 1


def bench_batch(vx, args) -> dict:
    jobs = batch_jobs(4, 256, 10, 18, 8)
    if args.batch_classes:  # a subset of the size classes (diagnostic runs)
        lo, hi = (int(x) for x in args.batch_classes.split(","))
        jobs = [j for j in jobs if lo <= j[0] <= hi]
    n_proofs = sum(len(ab) for _, ab in jobs)
    rows = sum((1 << l) * len(ab) for l, ab in jobs)
    nq, last = 80, 3
    # one job at a time on the plain path: one prover per size, each proof to host bytes
    # (--no-serial skips it, for profiles of the batched pass alone)
    solo_out, t_solo = None, None
    if not args.no_serial:
        solo = {l: vx.Prover(l, nq, last) for l in sorted({l for l, _ in jobs})}
        for l, p in solo.items():
            p.prove(1, 1)
        t = time.perf_counter()
        solo_out = []
        for l, ab in jobs:
            solo_out.append([solo[l].prove(a, b) for a, b in ab])
        t_solo = time.perf_counter() - t
        for p in solo.values():
            p.close()
        log(f"batch: {n_proofs} proofs one at a time: {t_solo:.3f}s ({n_proofs / t_solo:.1f} proofs/s)")
    # batched: all jobs submitted at once; provers made ahead (vx_server_prewarm, two per
    # size class) and graphs captured by a first pass, so the timed pass sets nothing up;
    # big batches staggered (the whole list arrives at once, so they would otherwise all
    # reach their FRI chains and proof copies together at the end)
    srv = vx.Server(0, nq, last, result_bytes=4 << 30)
    srv.set_graphs(True)
    srv.set_order(args.batch_order == "smallest")
    srv.set_big_batches(args.big_batches, 23)
    srv.set_stagger(not args.no_stagger)
    srv.set_copy_kernel(args.copy_kernel)
    for l in sorted({l for l, _ in jobs}):
        srv.prewarm(l, 2)
    warm = [srv.submit(l, ab) for l, ab in jobs]
    srv.wait_all()
    for i in warm:
        srv.release(i)
    # timed passes (--batch-reps): the median is reported, every pass's proofs checked
    passes = []
    same_all = True
    for rep in range(max(1, args.batch_reps)):
        t = time.perf_counter()
        ids = [srv.submit(l, ab) for l, ab in jobs]
        t_submit = time.perf_counter() - t
        srv.poll()  # the first scheduling step launches a batch on every idle prover
        t_first = time.perf_counter() - t - t_submit
        srv.wait_all()
        passes.append((time.perf_counter() - t, t_submit, t_first, ids))
        if rep + 1 < max(1, args.batch_reps):
            if solo_out is not None:
                same_all = same_all and all(srv.proof(i, c, l) == solo_out[k][c]
                                            for k, (i, (l, ab)) in enumerate(zip(ids, jobs)) for c in range(len(ab)))
            for i in ids:
                srv.release(i)
    all_s = sorted(p[0] for p in passes)
    last_ids = passes[-1][3]  # the last pass's proofs are still held by the server
    t_batch, t_submit, t_first, ids = sorted(passes, key=lambda p: p[0])[len(passes) // 2]
    log(f"batch: passes {[round(x, 4) for x in all_s]} s; host time to submit {1e3 * t_submit:.2f} ms, "
        f"first scheduling step {1e3 * t_first:.2f} ms")
    lat = [srv.latency_ms(i) for i in ids]
    by_class = {}
    for (l, _), x in zip(jobs, lat):
        by_class.setdefault(l, []).append(x)
    class_done = {str(l): round(max(v), 2) for l, v in sorted(by_class.items())}
    log(f"batch: last job of each size class done at (ms) {class_done}")
    same = None if solo_out is None else same_all and all(
        srv.proof(i, c, l) == solo_out[k][c] for k, (i, (l, ab)) in enumerate(zip(last_ids, jobs))
        for c in range(len(ab)))
    st = srv.stats()
    srv.close()
    log(f"batch: batched {t_batch:.3f}s ({n_proofs / t_batch:.1f} proofs/s), "
        f"{(t_solo or 0) / t_batch:.2f}x, identical to one-at-a-time: {same}")
    # the VALU bound of the job list: its Blake2s compressions at the VALU peak
    bound_s = sum(proof_compressions(l, last) * len(ab) for l, ab in jobs) * OPS_PER_COMPRESSION / (VALU_PEAK_TOPS * 1e12)
    cpu = None
    if not args.no_cpu:
        import oracle_lib

        threads = HOST_THREADS
        oracle_lib.lib().orc_set_threads(threads)
        sample = [(l, a, b) for l, ab in jobs[:12] for a, b in ab][:32]
        t = time.perf_counter()
        for l, a, b in sample:
            oracle_lib.prove(a, b, l, nq, last)
        secs = time.perf_counter() - t
        cpu = {"value": round(len(sample) / secs, 3), "unit": "proofs/s", "cores": threads, "kind": "port",
               "sample": f"oracle proofs of the first {len(sample)} proofs of the same job list, {secs:.2f}s"}
    return {"workload": "256 jobs of 1-8 Fibonacci columns, log_n uniform in [10,18] (splitmix64 seed 4), "
                        "blowup 4, 80 queries, last layer 2^3; every column one KRAK proof",
            "n_proofs": n_proofs, "rows": rows, "batched_proofs_per_s": round(n_proofs / t_batch, 2),
            "one_at_a_time_proofs_per_s": round(n_proofs / t_solo, 2) if t_solo else None,
            "speedup_vs_one_at_a_time": round(t_solo / t_batch, 3) if t_solo else None,
            "batched_s": round(t_batch, 4), "one_at_a_time_s": round(t_solo, 4) if t_solo else None,
            "batched_melem_s": round(rows / t_batch / 1e6, 2),
            "host_submit_ms": round(1e3 * t_submit, 3), "host_first_step_ms": round(1e3 * t_first, 3),
            "batched_passes_s": [round(x, 4) for x in all_s], "class_done_ms": class_done,
            "valu_bound_s": round(bound_s, 4), "batched_frac_of_valu_bound": round(bound_s / t_batch, 4),
            "job_latency_ms": quantiles(lat), "identical_to_one_at_a_time": same,
            "batches": st["batches"], "peak_device_bytes": st["peak_device_bytes"],
            "scheduling": {"order": args.batch_order, "staggered_big_batches": not args.no_stagger,
                           "proof_copies": "kernel" if args.copy_kernel else "dma", "graphs": True},
            "cpu_baseline": cpu}


def proof_compressions(log_n: int, log_last: int) -> int:
    """Blake2s compressions of one proof's trees: trace and quotient trees of E
    leaves (2E - 1 each) and one tree per FRI layer of E/2, E/4, ... values down to
    2^(log_last + 1) (2n - 1 each); transcript hashes are negligible."""
    E = 1 << (log_n + 2)
    c = 2 * (2 * E - 1)
    for j in range(log_n + 2 - log_last):
        n = E >> (j + 1)
        c += 2 * n - 1
    return c


def bench_serve(vx, args) -> dict:
    """The job list one at a time (a capacity-1 prover per size, every proof
    copied to the host and hashed), then two replays of the same list through
    the server: at the offered rate (latency) and with every job arriving
    within a fraction of a second (throughput at saturation). Every served
    proof is compared with its one-at-a-time bytes."""
    import inputs

    jobs = batch_jobs(5, args.serve_jobs, 12, 26, 34)
    u = (inputs.splitmix64(6, len(jobs)) >> np.uint64(11)).astype(np.float64) / float(1 << 53)
    gaps = -np.log1p(-u)
    nq, last = 80, 3
    n_proofs = sum(len(ab) for _, ab in jobs)
    rows = sum((1 << l) * len(ab) for l, ab in jobs)
    # one at a time: the whole job list in arrival order, each proof to host bytes
    sizes = sorted({l for l, _ in jobs})
    solo = {}
    for l in sizes:
        solo[l] = vx.Prover(l, nq, last)
        solo[l].prove(1, 1)
    t = time.perf_counter()
    digests = []
    for l, ab in jobs:
        digests.append([hashlib.sha256(solo[l].prove(a, b)).digest() for a, b in ab])
    t_serial = time.perf_counter() - t
    for p in solo.values():
        p.close()
    log(f"serve: {n_proofs} proofs one at a time: {t_serial:.3f}s")
    srv = vx.Server(288 << 30, nq, last, result_bytes=8 << 30)
    srv.set_stagger(args.serve_stagger)
    srv.set_copy_kernel(args.copy_kernel)
    # warm-up: one single-column job per size creates the provers (outside the replays)
    for l in sizes:
        srv.submit(l, [(1, 1)])
    srv.wait_all()
    for i in range(srv.stats()["jobs_done"]):
        srv.release(i)
    runs = {}
    identical = True
    for name, rate in (("offered", args.serve_rate), ("saturated", args.serve_rate * 50)):
        arrivals = np.cumsum(gaps / rate)
        arrivals -= arrivals[0]
        first = srv.stats()["jobs_done"]
        t = time.perf_counter()
        lat = srv.run_trace(arrivals, [l for l, _ in jobs], [len(ab) for _, ab in jobs],
                            [v for _, ab in jobs for pair in ab for v in pair])
        span = time.perf_counter() - t
        ok = all(srv.status(first + i) == 1 for i in range(len(jobs)))
        same = ok and all(hashlib.sha256(srv.proof(first + i, c, l)).digest() == digests[i][c]
                          for i, (l, ab) in enumerate(jobs) for c in range(len(ab)))
        identical = identical and same
        for i in range(len(jobs)):
            srv.release(first + i)
        runs[name] = {"rate_jobs_per_s": rate, "makespan_s": round(span, 4), "proofs_per_s": round(n_proofs / span, 2),
                      "melem_s": round(rows / span / 1e6, 2), "latency_ms": quantiles(list(lat)), "all_done": ok,
                      "identical_to_one_at_a_time": same}
        log(f"serve ({name}, {rate:.0f} jobs/s): {n_proofs} proofs in {span:.3f}s, {rows / span / 1e6:.0f} Melem/s, "
            f"p50 {np.percentile(lat, 50):.1f} ms p99 {np.percentile(lat, 99):.1f} ms, same bytes: {same}")
    st = srv.stats()
    srv.close()
    sat = runs["saturated"]["makespan_s"]
    # the VALU bound of the job list: its Blake2s compressions at the VALU peak
    comp = sum(proof_compressions(l, last) * len(ab) for l, ab in jobs)
    bound_s = comp * OPS_PER_COMPRESSION / (VALU_PEAK_TOPS * 1e12)
    log(f"serve: one at a time {t_serial:.3f}s -> {t_serial / sat:.2f}x at saturation; VALU bound {bound_s:.3f}s "
        f"(saturated at {bound_s / sat:.2f}, one at a time at {bound_s / t_serial:.2f} of it)")
    return {"workload": f"{len(jobs)} jobs, Poisson arrivals, 1-34 Fibonacci columns, log_n uniform in [12,26] "
                        "(splitmix64 seeds 5, 6), 80 queries, last layer 2^3, HBM budget 288 GB",
            "n_proofs": n_proofs, "rows": rows, **runs,
            "one_at_a_time_s": round(t_serial, 4),
            "one_at_a_time": "measured: the same job list proved in arrival order by a capacity-1 prover per size, "
                             "each proof copied to the host",
            "speedup_vs_one_at_a_time": round(t_serial / sat, 3),
            "identical_to_one_at_a_time": identical,
            "valu_bound_s": round(bound_s, 4), "compressions": comp,
            "saturated_frac_of_valu_bound": round(bound_s / sat, 4),
            "one_at_a_time_frac_of_valu_bound": round(bound_s / t_serial, 4),
            "max_speedup_at_valu_bound": round(t_serial / bound_s, 3),
            "peak_device_bytes": st["peak_device_bytes"], "batches": st["batches"]}


def bench_deep(vx, args) -> dict:
    """stwo DEEP quotient of the wide commitment's LDE columns (4 x 2^16, 10 x 2^20,
    20 x 2^22, splitmix64 seed 8) at two sample points: a batch per (point, size),
    six cuda_accumulate_numerators launches, then cuda_compute_quotients_combine
    over the 2^22-row lifting domain. HIP events on the null stream time each part."""
    import inputs
    import oracle_lib as O

    sizes = [16] * 4 + [20] * 10 + [22] * 20
    L = 22
    cols = inputs.m31_columns(8, sizes)
    rng = np.random.default_rng(8)
    q = lambda n: rng.integers(0, vx_p(), n, dtype=np.uint64).astype(np.uint32)
    dcols = [vx.DeviceBuffer.from_host(c) for c in cols]
    keep = []

    def dev(a):
        b = vx.DeviceBuffer.from_host(np.ascontiguousarray(a, dtype=np.uint32))
        keep.append(b)
        return b

    def dev_ptrs(bufs):
        addrs = np.array([b.addr for b in bufs], dtype=np.uint64)
        return dev(addrs.view(np.uint32))

    pts = [(q(4), q(4)), (q(4), q(4))]
    lc = q(2 * len(sizes) * 12).reshape(2, len(sizes), 12)  # a, b, c of column c at point p
    batches = []  # (log size, column indices, b, c, point)
    for pt in range(2):
        for lg in (16, 20, 22):
            idx = np.array([i for i, l in enumerate(sizes) if l == lg], np.uint32)
            batches.append((lg, idx, lc[pt, idx, 4:8].ravel().copy(), lc[pt, idx, 8:12].ravel().copy(), pt))
    d_colptrs = dev_ptrs(dcols)
    numer = [[vx.DeviceBuffer(1 << lg) for _ in range(4)] for lg, *_ in batches]
    d_args = [(dev(idx), dev(b), dev(c)) for _, idx, b, c, _ in batches]
    spx = np.concatenate([pts[pt][0] for *_, pt in batches])
    spy = np.concatenate([pts[pt][1] for *_, pt in batches])
    fla = np.concatenate([(lc[pt, idx, 0:4].astype(np.uint64).sum(axis=0) % vx_p()).astype(np.uint32)
                          for _, idx, _, _, pt in batches])
    logs = np.array([lg for lg, *_ in batches], np.uint32)
    dxs, dys = O.circle_domain_brev(L)  # the domain points are an input of the ABI
    d_spx, d_spy, d_fla, d_logs, d_dx, d_dy = (dev(x) for x in (spx, spy, fla, logs, dxs, dys))
    # the fused form on the same inputs (its domain points built on the device)
    fx, fy = vx.circle_domain_points(L)
    fused_out = [vx.DeviceBuffer(1 << L) for _ in range(4)]
    dq = vx.DeepQuotient(len(sizes), 2)
    points = np.stack([np.concatenate(p) for p in pts])

    def fused():
        dq.quotient([c.addr for c in dcols], sizes, points, lc, L, fx.addr, fy.addr, [o.addr for o in fused_out])
    d_np = [dev_ptrs([numer[j][k] for j in range(len(batches))]) for k in range(4)]
    out = [vx.DeviceBuffer(1 << L) for _ in range(4)]
    e = [vx.Event() for _ in range(3)]

    def step(timed):
        if timed:
            e[0].record()
        for (lg, idx, *_), (di, db, dc), nb in zip(batches, d_args, numer):
            vx.lib.cuda_accumulate_numerators(d_colptrs.ptr, di.ptr, db.ptr, dc.ptr, len(idx), 1 << lg,
                                              *[x.ptr for x in nb])
        if timed:
            e[1].record()
        vx.lib.cuda_compute_quotients_combine(d_spx.ptr, d_spy.ptr, d_fla.ptr, *[p.ptr for p in d_np], d_logs.ptr,
                                              len(batches), d_dx.ptr, d_dy.ptr, L, 1 << L, *[o.ptr for o in out])
        if timed:
            e[2].record()
            return e[0].elapsed_ms(e[1]), e[1].elapsed_ms(e[2])
        return None

    for _ in range(3):
        step(False)
    vx.synchronize()
    vx.check("deep")
    K = 20
    t_num = t_comb = 0.0
    for _ in range(K):
        a, b = step(True)
        t_num += a
        t_comb += b
    t_num /= K
    t_comb /= K
    for _ in range(3):
        fused()
    # K calls back to back between two events: each call's host-side metadata
    # (two staging slots) is prepared while the previous call's kernel runs
    e[0].record()
    for _ in range(K):
        fused()
    e[1].record()
    t_fused = e[0].elapsed_ms(e[1]) / K
    rows = 1 << L
    num_bytes = sum(4 * len(idx) * (1 << int(lg)) + 16 * (1 << int(lg)) for lg, idx, *_ in batches)
    comb_bytes = rows * (8 + 16) + sum(16 * (1 << int(lg)) for lg in logs)
    bw = num_bytes / (t_num * 1e-3) / 1e9
    ops = OPS_PER_ROW_ACC * rows * len(batches) / (t_comb * 1e-3) / 1e12
    # parity: the combined quotient against the oracle on the same inputs
    got = np.stack([o.to_host() for o in out])
    t = time.perf_counter()
    exp_num = [O.accumulate_numerators(cols, idx, b, c, 1 << lg) for lg, idx, b, c, _ in batches]
    exp = O.quotients_combine(spx, spy, fla, exp_num, logs, dxs, dys, L)
    cpu_s = time.perf_counter() - t
    parity = bool(np.array_equal(got, exp))
    fused_same = bool(np.array_equal(np.stack([o.to_host() for o in fused_out]), got))
    dq.close()
    ms = t_num + t_comb
    # fused: each column read once, one denominator per (row, point)
    fused_bytes = sum(4 * (1 << l) for l in sizes) + rows * (8 + 16)
    fused_ops = FUSED_OPS_PER_ROW * rows
    log(f"deep: numerators {t_num:.3f} ms ({bw:.0f} GB/s), combine {t_comb:.3f} ms ({ops:.2f} Tops/s), "
        f"{rows / (ms * 1e-3) / 1e6:.0f} Mrow/s, parity={parity}; fused {t_fused:.3f} ms "
        f"({fused_ops / (t_fused * 1e-3) / 1e12:.2f} Tops/s), same={fused_same}; oracle {cpu_s:.2f} s")
    res = {"workload": "stwo DEEP quotient: 34 M31 LDE columns (4x2^16, 10x2^20, 20x2^22, splitmix64 seed 8), "
                       "2 sample points, 6 accumulations (one per point and size), combined on the 2^22 "
                       "lifting domain; reference ABI kernels, inputs resident in HBM",
           "ms_per_quotient": round(ms, 4), "mrows_per_s": round(rows / (ms * 1e-3) / 1e6, 2),
           "numerators_ms": round(t_num, 4), "combine_ms": round(t_comb, 4),
           "roofline_numerators": {"bound": "hbm", "kernel": "accumulate_numerators_kernel (6 launches)",
                                   "achieved": round(bw, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                   "frac": round(bw / HBM_PEAK_GBS, 4), "bytes": num_bytes,
                                   "traffic": TRAFFIC.get("deep:numerators")},
           "roofline_combine": {"bound": "valu", "kernel": "quotients_combine_kernel",
                                "achieved": round(ops, 3), "peak": round(VALU_PEAK_TOPS, 2),
                                "unit": "Tops/s (int32 lane-ops)", "frac": round(ops / VALU_PEAK_TOPS, 4),
                                "ops_per_row_acc": OPS_PER_ROW_ACC, "hbm_bytes": comb_bytes,
                                "traffic": TRAFFIC.get("deep:combine")},
           "parity_vs_oracle": parity,
           "fused": {"kernel": "deep_fused_kernel<2,4> (vx_deep_quotient: one launch, columns read once, "
                               "one denominator per row and point)",
                     "ms_per_quotient": round(t_fused, 4), "mrows_per_s": round(rows / (t_fused * 1e-3) / 1e6, 2),
                     "speedup_vs_reference_pipeline": round(ms / t_fused, 3), "same_words_as_abi": fused_same,
                     "roofline": {"bound": "valu", "achieved": round(fused_ops / (t_fused * 1e-3) / 1e12, 3),
                                  "peak": round(VALU_PEAK_TOPS, 2), "unit": "Tops/s (int32 lane-ops)",
                                  "frac": round(fused_ops / (t_fused * 1e-3) / 1e12 / VALU_PEAK_TOPS, 4),
                                  "ops_per_row": FUSED_OPS_PER_ROW},
                     "roofline_hbm": {"bound": "hbm", "achieved": round(fused_bytes / (t_fused * 1e-3) / 1e9, 1),
                                      "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                      "frac": round(fused_bytes / (t_fused * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                                      "bytes": fused_bytes, "traffic": TRAFFIC.get("deep:fused")}}}
    if not args.no_cpu:
        res["cpu_baseline"] = {"value": round(rows / cpu_s / 1e6, 3), "unit": "Mrow/s",
                               "cores": int(O.lib().orc_get_threads()), "kind": "port",
                               "sample": "the whole workload (oracle accumulate_numerators + quotients_combine)"}
    return res


def bench_host(vx, args) -> dict:
    """The plain workload with the trace handed over in host memory (the
    reference's flow: trace in a pinned buffer, H2D, prove): vx_prover_prove_batch_host
    calls back to back, so each call's 2^log_n-word upload overlaps the previous
    proof. Pinned (vx_malloc_host) and pageable (numpy) traces. PCIe-inclusive:
    reported beside the headline, never as it."""
    log_n, nq, last = args.log_n, args.queries, args.log_last
    a, b = 1262994060, 193768037
    prover = vx.Prover(log_n, nq, last)
    d = vx.fibonacci_trace_device(a, b, log_n)
    vx.synchronize()
    pinned = vx.PinnedBuffer(1 << log_n)
    pinned.array[:] = d.to_host()
    pageable = pinned.array.copy()
    out = {}
    for name, trace in (("pinned", pinned.addr), ("pageable", pageable)):
        for _ in range(2):
            prover.prove_batch_host([trace], [(a, b)])
        vx.synchronize_stream(prover.stream)
        K = max(4, args.steps // 2)
        t = time.perf_counter()
        for _ in range(K):
            prover.prove_batch_host([trace], [(a, b)])
        vx.synchronize_stream(prover.stream)
        ms = (time.perf_counter() - t) / K * 1e3
        out[name] = {"ms_per_proof": round(ms, 4), "melem_s": round((1 << log_n) / (ms * 1e-3) / 1e6, 2)}
    same = prover.proof_bytes(0) == vx.Prover(log_n, nq, last).prove(a, b)
    prover.close()
    log(f"host traces: pinned {out['pinned']['ms_per_proof']:.3f} ms/proof, pageable "
        f"{out['pageable']['ms_per_proof']:.3f} ms/proof, same proof: {same}")
    return {"workload": "plain workload (2^%d rows) with the trace in host memory, uploaded per proof on the "
                        "prover's upload stream while the previous proof runs" % log_n,
            **out, "trace_bytes": 4 << log_n, "same_proof_as_device_trace": same}


def bench_grind(vx, args) -> dict:
    """Blake2s PoW grinding at the reference's 26 bits: K channel states (splitmix64
    seed 7), each grind one vx_grind_blake2s call (prefix on the device, search,
    nonce in HBM), K calls enqueued back to back on the null stream and timed
    between synchronisations. The work is the nonces a search must hash, found + 1
    per state; VALU roofline at OPS_PER_NONCE."""
    import inputs
    import oracle_lib as O

    K = 256
    seeds = inputs.splitmix64(7, 8 * K).astype(np.uint32).reshape(K, 8)
    chans = [vx.DeviceBuffer.from_host(np.concatenate([st, np.zeros(8, np.uint32)])) for st in seeds]
    res = vx.DeviceBuffer(2 * K)

    def run(n):
        for i in range(n):
            vx.grind_device(chans[i].addr, POW_BITS, res.addr + 8 * i)
        vx.synchronize()

    run(8)  # warm-up
    best = None
    for _ in range(3):
        t = time.perf_counter()
        run(K)
        dt = time.perf_counter() - t
        best = dt if best is None else min(best, dt)
    words = res.to_host(2 * K).astype(np.uint64)
    nonces = [int(x) for x in (words[0::2] | (words[1::2] << np.uint64(32)))]
    tried = sum(n + 1 for n in nonces)
    rate = tried / best
    achieved = rate * OPS_PER_NONCE / 1e12
    # parity: every nonce valid; the first two are the smallest valid ones (oracle search)
    valid = all(O.pow_zeros(O.pow_prefix(seeds[i], POW_BITS), nonces[i]) >= POW_BITS for i in range(K))
    first = all(O.pow_search(O.pow_prefix(seeds[i], POW_BITS), POW_BITS, False, 0, nonces[i] + 1) == nonces[i]
                for i in range(2))
    log(f"grind: {K} x {POW_BITS}-bit PoW in {best * 1e3:.1f} ms ({best / K * 1e3:.3f} ms each), "
        f"{rate / 1e9:.2f} G nonces/s, valid={valid} first={first}")
    # the reference's host loop over the ABI (grind_ops.rs:70-115): prefix on the host,
    # batches of 2^24 nonces, a synchronisation and a result read after each batch
    R = 32
    prefixes = [O.pow_prefix(seeds[i], POW_BITS) for i in range(R)]
    d_pre = vx.DeviceBuffer(8)
    d_res = vx.DeviceBuffer(2)
    ref_nonces = []
    t = time.perf_counter()
    for i in range(R):
        vx.lib.vx_memcpy_h2d(d_pre.ptr, prefixes[i].ctypes.data, 32)
        vx.lib.vx_memcpy_h2d(d_res.ptr, np.array([~np.uint64(0)], np.uint64).ctypes.data, 8)
        off = 0
        while True:
            vx.lib.cuda_grind_pow(d_pre.ptr, d_res.ptr, POW_BITS, off, 1 << 24)
            vx.synchronize()
            lo, hi = d_res.to_host(2)
            n = int(lo) | (int(hi) << 32)
            if n != (1 << 64) - 1:
                break
            off += 1 << 24
        ref_nonces.append(n)
    t_ref = time.perf_counter() - t
    same_as_device = ref_nonces == nonces[:R]
    ref_tried = sum(n + 1 for n in ref_nonces)
    log(f"grind, reference host loop (2^24-nonce batches): {t_ref / R * 1e3:.3f} ms each, "
        f"{ref_tried / t_ref / 1e9:.2f} G nonces/s, same nonces: {same_as_device}")
    out = {"workload": f"{K} Blake2s PoW grinds at pow_bits {POW_BITS} (channel states splitmix64 seed 7), "
                       "prefix + search + nonce on the device, one call each",
           "ms_per_grind": round(best / K * 1e3, 4), "nonces_per_grind_mean": round(tried / K),
           "g_nonces_per_s": round(rate / 1e9, 3),
           "roofline": {"bound": "valu", "kernel": "grind_kernel<false,false,true>", "achieved": round(achieved, 3),
                        "peak": round(VALU_PEAK_TOPS, 2), "unit": "Tops/s (int32 lane-ops)",
                        "frac": round(achieved / VALU_PEAK_TOPS, 4), "ops_per_nonce": OPS_PER_NONCE,
                        "peak_measured": round(MEASURED.get("valu_peak", 0.0), 3)},
           "parity_vs_oracle": {"all_valid": valid, "first_two_smallest": first},
           "reference_host_loop": {"what": "cuda_grind_pow in batches of 2^24 with a host synchronisation and "
                                           "result read per batch, as grind_ops.rs:70-115, first 32 states",
                                   "ms_per_grind": round(t_ref / R * 1e3, 4),
                                   "g_nonces_per_s": round(ref_tried / t_ref / 1e9, 3),
                                   "same_nonces_as_device_step": same_as_device}}
    if not args.no_cpu:
        threads = int(O.lib().orc_get_threads())
        sample = 1 << 27
        pre = O.pow_prefix(seeds[0], POW_BITS)
        t = time.perf_counter()
        O.pow_search(pre, 129, False, 0, sample)  # unreachable target: every nonce hashed
        secs = time.perf_counter() - t
        out["cpu_baseline"] = {"value": round(sample / secs / 1e9, 5), "unit": "G nonces/s", "cores": threads,
                               "kind": "port", "sample": "2^27 nonces, target never met (oracle orc_pow_search)"}
        log(f"grind cpu: {sample / secs / 1e6:.1f} M nonces/s on {threads} threads")
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--log-n", type=int, default=24)
    ap.add_argument("--queries", type=int, default=100)
    ap.add_argument("--log-last", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--no-pipelined", action="store_true", help="skip the multi-stream plain measurement")
    ap.add_argument("--batch-reps", type=int, default=3, help="timed passes of the batch workload (median reported)")
    ap.add_argument("--big-batches", type=int, default=0,
                    help="batches of >= 2^23 rows in flight at once in the batch workload (0: no limit)")
    ap.add_argument("--no-stagger", action="store_true",
                    help="big batches of the batch workload (all jobs submitted at once) start at once instead "
                         "of after the heavy part of the big batch before them")
    ap.add_argument("--serve-stagger", action="store_true",
                    help="stagger the big batches of the serve workload too (off: arrivals already spread them)")
    ap.add_argument("--copy-kernel", action="store_true",
                    help="the batch and serve workloads copy proofs with a kernel instead of the DMA engines")
    ap.add_argument("--batch-classes", default="", help="lo,hi: only the jobs of log_n in [lo, hi] (diagnostics)")
    ap.add_argument("--batch-order", choices=["smallest", "largest"], default="smallest",
                    help="launch order of size classes in the server's scheduling steps")
    ap.add_argument("--no-serial", action="store_true",
                    help="skip the one-at-a-time pass of the batch workload (profiles of the batched pass alone)")
    ap.add_argument("--cpu-log-n", type=int, default=24)
    ap.add_argument("--cpu-reps", type=int, default=3)
    ap.add_argument("--skip", default="", help="comma list of workloads to skip: plain,wide,batch,serve,host,deep,grind "
                                               "(skipping plain leaves value null: for profiling runs)")
    ap.add_argument("--serve-jobs", type=int, default=48)
    ap.add_argument("--serve-rate", type=float, default=4.0, help="offered load, jobs/s")
    ap.add_argument("--no-tune", action="store_true", help="keep the default NTT pass split")
    ap.add_argument("--lde-plan", action="append", default=[], metavar="K:m,c,r/m,c,r...",
                    help="pin the blowup-4 evaluation plan of 2^K points after tuning (repeatable), e.g. "
                         "26:12,0,0/6,6,0/8,6,1 (counter runs that must see one plan)")
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="also run a device-to-device copy of this many MiB (counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path here "
              "(independent proofs would shard by job). Run with --gpus 1.", file=sys.stderr)
        return 2

    import vortexstark_amd as vx

    assert vx.gpu_available(), "no HIP device"
    skip = {x for x in args.skip.split(",") if x}
    MEASURED["valu_peak"] = vx.measure_valu_peak()
    log(f"measured int32 VALU peak: {MEASURED['valu_peak']:.2f} T lane-ops/s (nominal {VALU_PEAK_TOPS:.2f})")
    MEASURED["hash_latency_ns"] = dict(zip(("quad", "lane"), (round(x, 1) for x in vx.measure_hash_latency())))
    log(f"Merkle node hash latency on one wave: {MEASURED['hash_latency_ns']['quad']:.0f} ns quad-cooperative, "
        f"{MEASURED['hash_latency_ns']['lane']:.0f} ns one lane")
    if not args.no_tune:
        t = time.time()
        MEASURED["ntt_plans"] = {str(k): {"ms": round(ms, 4), "passes": plan, "lde_passes": vx.ntt_lde_plan(k)}
                                 for k, (ms, plan) in vx.ntt_autotune(10, 28).items()}
        log(f"ntt autotune 2^10..2^28: {time.time() - t:.1f}s; blowup-4 evaluation plans (layers, segment, "
            f"replicating): " + ", ".join(f"2^{k}: {v['lde_passes']}" for k, v in MEASURED["ntt_plans"].items()
                                          if int(k) in (16, 20, 22, 26)))
        t = time.time()
        MEASURED["merkle_stage_plans"] = {str(k): {"ms": round(ms, 4), "logc": c}
                                          for k, (ms, c) in vx.merkle_autotune(12, 27).items()}
        log(f"merkle stage autotune 2^12..2^27: {time.time() - t:.1f}s "
            f"{ {k: v['logc'] for k, v in MEASURED['merkle_stage_plans'].items()} }")
    for spec in args.lde_plan:
        k, passes = spec.split(":")
        plan = [tuple(int(x) for x in p.split(",")) for p in passes.split("/")]
        vx.ntt_lde_plan_set(int(k), plan)
        log(f"blowup-4 plan of 2^{k} pinned: {vx.ntt_lde_plan(int(k))}")
    extra = {}
    if "plain" in skip:  # profiling runs of the other workloads only
        calibration_copy(vx, args)
        for name, fn in (("wide", bench_wide), ("batch", bench_batch), ("serve", bench_serve),
                         ("host", bench_host), ("deep", bench_deep), ("grind", bench_grind)):
            if name not in skip:
                extra["host_trace" if name == "host" else name] = fn(vx, args)
        print(json.dumps({"metric": METRIC, "value": None, "unit": "Melem/s", "n_gpus": 1, **extra}), flush=True)
        return 0
    plain = bench_plain(vx, args)
    if "wide" not in skip:
        extra["wide"] = bench_wide(vx, args)
    if "batch" not in skip:
        extra["batch"] = bench_batch(vx, args)
    if "serve" not in skip:
        extra["serve"] = bench_serve(vx, args)
    if "host" not in skip:
        extra["host_trace"] = bench_host(vx, args)
    if "deep" not in skip:
        extra["deep"] = bench_deep(vx, args)
    if "grind" not in skip:
        extra["grind"] = bench_grind(vx, args)
    log_n, nq, last = args.log_n, args.queries, args.log_last
    out = {
        "metric": METRIC,
        "value": round(plain["melem_s"], 2),
        "unit": "Melem/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(plain["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32 (exact M31/QM31)",
        "data": "synthetic",
        "config": {"workload": "plain: Fibonacci M31 column 2^%d rows, blowup 4 (2^%d LDE), Blake2s Merkle, "
                               "FRI to 2^%d, %d queries, full proof in HBM" % (log_n, log_n + 2, last, nq),
                   "log_n": log_n, "blowup": 4, "n_queries": nq, "log_last_layer": last, "parallelism": "none"},
        "proofs_per_s": round(plain["proofs_per_s"], 3),
        "latency_ms": plain["latency"],
        "roofline": plain["roofline"],
        "roofline_hbm": plain["roofline_hbm"],
        "roofline_ntt_valu": plain["roofline_ntt"],
        "pipelined": plain["pipelined"],
        "pipeline_valu_bound_ms": round(plain["t_valu_ms"], 3),
        "pipeline_frac_of_valu_bound": round(plain["t_valu_ms"] / plain["ms_per_step"], 4),
        "phases_ms": {k: round(v, 4) for k, v in plain["phases"].items()},
        "parity_vs_oracle": plain["parity"],
        "cpu_baseline": plain["cpu"],
        "valu_peak_measured": round(MEASURED["valu_peak"], 3),
        "hash_latency_ns": MEASURED.get("hash_latency_ns"),
        "hip_hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4")),
        "merkle_stage_plans": MEASURED.get("merkle_stage_plans"),
        "ntt_plans": MEASURED.get("ntt_plans"),
    }
    out.update(extra)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
