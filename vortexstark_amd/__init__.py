"""vortexstark_amd — MI355X-native commit pipeline for VortexSTARK's prover.

Host-side mirror of the reference's prover interface (src/prover.rs:192-231):

    prove(a, b, log_n)              -> KRAK v1 proof bytes  (prove + serialize_proof)
    Prover(log_n).prove(a, b)       -> same, with the ProverCache reused
    WideCommit(log_sizes)           -> batched stwo-basis LDE + lifted tree + FRI commitments
    Server(budget)                  -> continuous batching of proof jobs under an HBM budget
    DeviceBuffer                    -> device memory owned from Python

Everything runs in libvortex_hip.so (HIP kernels for gfx950 behind a C ABI).
Importing this package fails if that library has not been built.
"""

from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _ffi

lib = _ffi.load()

P = 0x7FFFFFFF
BLOWUP_BITS = 2  # prover.rs:86
N_QUERIES = 80  # prover.rs:90
LOG_LAST_LAYER = 3  # fri.rs:265


class HipError(RuntimeError):
    pass


def check(what: str = "") -> None:
    code = lib.vx_get_last_error()
    if code:
        msg = lib.vx_get_last_error_message().decode(errors="replace")
        raise HipError(f"{what}: error {code}: {msg}")


def gpu_available() -> bool:
    return lib.vx_device_count() > 0


class DeviceBuffer:
    """Device allocation of n u32 words (src/device/buffer.rs: DeviceBuffer<u32>)."""

    def __init__(self, n_words: int):
        self.n = int(n_words)
        self.ptr = C.c_void_p()
        if lib.vx_malloc(C.byref(self.ptr), max(4, 4 * self.n)):
            check("vx_malloc")

    @classmethod
    def from_host(cls, arr) -> "DeviceBuffer":
        a = np.ascontiguousarray(arr, dtype=np.uint32)
        b = cls(a.size)
        if a.size and lib.vx_memcpy_h2d(b.ptr, a.ctypes.data, 4 * a.size):
            check("vx_memcpy_h2d")
        return b

    def to_host(self, n: Optional[int] = None) -> np.ndarray:
        n = self.n if n is None else n
        out = np.empty(n, dtype=np.uint32)
        if n and lib.vx_memcpy_d2h(out.ctypes.data, self.ptr, 4 * n):
            check("vx_memcpy_d2h")
        return out

    @property
    def addr(self) -> int:
        return self.ptr.value or 0

    def free(self) -> None:
        if self.ptr and self.ptr.value:
            lib.vx_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class PinnedBuffer:
    """Page-locked host memory of n u32 words (vx_malloc_host; the reference's
    PinnedBuffer<u32>, buffer.rs:472-532): DMA reads it while the host goes on.
    `array` is a numpy view of it."""

    def __init__(self, n_words: int):
        self.n = int(n_words)
        self.ptr = C.c_void_p()
        if lib.vx_malloc_host(C.byref(self.ptr), max(4, 4 * self.n)):
            check("vx_malloc_host")
            raise HipError("vx_malloc_host failed")
        self.array = np.ctypeslib.as_array((C.c_uint32 * max(1, self.n)).from_address(self.ptr.value))[: self.n]

    @property
    def addr(self) -> int:
        return self.ptr.value or 0

    def free(self) -> None:
        if self.ptr and self.ptr.value:
            self.array = None
            lib.vx_free_host(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_copy_words(addr: int, n: int) -> np.ndarray:
    out = np.empty(n, dtype=np.uint32)
    if n and lib.vx_memcpy_d2h(out.ctypes.data, C.c_void_p(addr), 4 * n):
        check("vx_memcpy_d2h")
    return out


def synchronize() -> None:
    if lib.vx_device_synchronize():
        check("vx_device_synchronize")


class Prover:
    """A prover for traces of 2^log_n rows; twiddles and workspace live on the device.

    Mirrors ProverCache + prove_cached (prover.rs:143-209); n_queries and
    log_last_layer default to the reference protocol (80 queries, last layer 2^3).
    """

    def __init__(self, log_n: int, n_queries: int = N_QUERIES, log_last_layer: int = LOG_LAST_LAYER,
                 capacity: int = 1):
        self.log_n = log_n
        self.n_queries = n_queries
        self.log_last_layer = log_last_layer
        self.capacity = capacity
        self.handle = lib.vx_prover_create_batch(log_n, n_queries, log_last_layer, capacity)
        if not self.handle:
            check("vx_prover_create")
            raise HipError("vx_prover_create failed")
        self.proof_size = int(lib.vx_prover_proof_size(self.handle))
        self.device_bytes = int(lib.vx_prover_device_bytes(self.handle))

    @property
    def stream(self) -> int:
        return lib.vx_prover_stream(self.handle) or 0

    def prove(self, a: int, b: int) -> bytes:
        """prove(a, b, log_n) + serialize_proof: KRAK v1 bytes."""
        buf = (C.c_uint8 * self.proof_size)()
        n = lib.vx_prove_fibonacci(self.handle, a % P, b % P, buf, self.proof_size)
        if n != self.proof_size:
            check("vx_prove_fibonacci")
            raise HipError("vx_prove_fibonacci failed")
        return bytes(buf)

    def prove_device(self, d_trace_addr: int, a: int, b: int) -> None:
        """Enqueue a proof of the trace already in device memory (no host wait)."""
        if lib.vx_prover_prove_device(self.handle, C.c_void_p(d_trace_addr), a, b):
            check("vx_prover_prove_device")

    def prove_batch(self, ab) -> list:
        """Fibonacci proofs of every (a, b) in one batch (traces generated on the device)."""
        self.prove_batch_async(ab)
        return [self.proof_bytes(i) for i in range(len(ab))]

    def prove_batch_async(self, ab) -> None:
        flat = (C.c_uint32 * (2 * len(ab)))(*[int(v) % P for pair in ab for v in pair])
        if lib.vx_prove_fibonacci_batch(self.handle, flat, len(ab)):
            check("vx_prove_fibonacci_batch")

    def prove_batch_device(self, trace_addrs, ab) -> None:
        """Enqueue proofs of traces already in device memory, one batch."""
        ptrs = (C.c_void_p * len(trace_addrs))(*trace_addrs)
        flat = (C.c_uint32 * (2 * len(ab)))(*[int(v) % P for pair in ab for v in pair])
        if lib.vx_prover_prove_batch_device(self.handle, ptrs, flat, len(trace_addrs)):
            check("vx_prover_prove_batch_device")

    def prove_batch_host(self, traces, ab) -> None:
        """Enqueue proofs of traces in host memory: numpy uint32 arrays (copied
        before this returns) or addresses of vx_malloc_host memory (read by DMA
        while the previous call proves)."""
        keep, addrs = [], []
        for t in traces:
            if isinstance(t, int):
                addrs.append(t)
            else:
                a = np.ascontiguousarray(t, dtype=np.uint32)
                keep.append(a)
                addrs.append(a.ctypes.data)
        ptrs = (C.c_void_p * len(addrs))(*addrs)
        flat = (C.c_uint32 * (2 * len(ab)))(*[int(v) % P for pair in ab for v in pair])
        if lib.vx_prover_prove_batch_host(self.handle, ptrs, flat, len(addrs)):
            check("vx_prover_prove_batch_host")
            raise HipError("vx_prover_prove_batch_host failed")

    def set_graph(self, on: bool) -> None:
        if lib.vx_prover_set_graph(self.handle, 1 if on else 0):
            check("vx_prover_set_graph")

    def proof_bytes(self, index: int = 0) -> bytes:
        buf = (C.c_uint8 * self.proof_size)()
        n = lib.vx_prover_proof_copy_at(self.handle, index, buf, self.proof_size)
        if n != self.proof_size:
            check("vx_prover_proof_copy")
            raise HipError("vx_prover_proof_copy failed")
        return bytes(buf)

    def eval_words(self) -> np.ndarray:
        lib.vx_stream_synchronize(C.c_void_p(self.stream))
        return device_copy_words(lib.vx_prover_eval_device(self.handle), 1 << (self.log_n + BLOWUP_BITS))

    def fri_layer(self, j: int) -> np.ndarray:
        """FRI layer j as a (4, size) array of SoA columns."""
        lib.vx_stream_synchronize(C.c_void_p(self.stream))
        size = 1 << (self.log_n + BLOWUP_BITS - 1 - j)
        cols = [device_copy_words(lib.vx_prover_fri_layer_device(self.handle, j, c), size) for c in range(4)]
        return np.stack(cols)

    def close(self) -> None:
        if self.handle:
            lib.vx_prover_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def prove(a: int, b: int, log_n: int, n_queries: int = N_QUERIES, log_last_layer: int = LOG_LAST_LAYER) -> bytes:
    """prover::prove(a, b, log_n) followed by serialize_proof (KRAK v1 bytes)."""
    p = Prover(log_n, n_queries, log_last_layer)
    try:
        return p.prove(a, b)
    finally:
        p.close()


def fibonacci_trace_device(a: int, b: int, log_n: int) -> DeviceBuffer:
    buf = DeviceBuffer(1 << log_n)
    if lib.vx_fibonacci_trace(a % P, b % P, log_n, buf.ptr, None):
        check("vx_fibonacci_trace")
    synchronize()
    return buf


def _timeline(count, label, ms, handle) -> dict:
    out = {}
    for i in range(1, count(handle)):
        key = label(handle, i).decode()
        out[key] = out.get(key, 0.0) + ms(handle, i)
    return out


class WideCommit:
    """Batched wide commitment over M31 columns of mixed sizes (Part 4 of the C ABI).

    Columns are committed in ascending log-size order, the order stwo gives the
    columns of one lifted tree (crates/cuda-backend/src/merkle_ops.rs:49).
    """

    def __init__(self, log_sizes, log_blowup: int = BLOWUP_BITS, log_last_layer: int = 4):
        self.log_sizes = [int(x) for x in log_sizes]
        arr = (C.c_uint32 * len(self.log_sizes))(*self.log_sizes)
        self.handle = lib.vx_wide_create(len(self.log_sizes), arr, log_blowup, log_last_layer)
        if not self.handle:
            check("vx_wide_create")
            raise HipError("vx_wide_create failed")
        self.log_blowup = log_blowup
        self.log_last_layer = log_last_layer
        self.n_fri = int(lib.vx_wide_n_fri_layers(self.handle))
        self.lifting_log_size = int(lib.vx_wide_lifting_log_size(self.handle))

    @property
    def stream(self) -> int:
        return lib.vx_wide_stream(self.handle) or 0

    def commit_device(self, col_addrs) -> None:
        ptrs = (C.c_void_p * len(col_addrs))(*col_addrs)
        if lib.vx_wide_commit_device(self.handle, ptrs):
            check("vx_wide_commit_device")

    def results(self):
        """(root[8], fri_roots[n_fri][8], last[2^log_last][4], beta[4]) as numpy arrays."""
        root = np.zeros(8, np.uint32)
        fri = np.zeros(8 * self.n_fri, np.uint32)
        last = np.zeros(4 << self.log_last_layer, np.uint32)
        beta = np.zeros(4, np.uint32)
        if lib.vx_wide_results(self.handle, root.ctypes.data, fri.ctypes.data, last.ctypes.data, beta.ctypes.data):
            check("vx_wide_results")
            raise HipError("vx_wide_results failed")
        return root, fri.reshape(-1, 8), last.reshape(-1, 4), beta

    def lde(self, c: int) -> np.ndarray:
        synchronize_stream(self.stream)
        n = 1 << (self.log_sizes[c] + self.log_blowup)
        return device_copy_words(lib.vx_wide_lde_device(self.handle, c), n)

    def set_timing(self, on: bool) -> None:
        lib.vx_wide_set_timing(self.handle, 1 if on else 0)

    def timeline(self) -> dict:
        return _timeline(lib.vx_wide_timeline_count, lib.vx_wide_timeline_label, lib.vx_wide_timeline_ms,
                         self.handle)

    def close(self) -> None:
        if self.handle:
            lib.vx_wide_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Event:
    """A HIP event (vx_event_*), for timing launches on a stream."""

    def __init__(self):
        self.ptr = C.c_void_p()
        if lib.vx_event_create(C.byref(self.ptr)):
            check("vx_event_create")

    def record(self, stream: int = 0) -> None:
        if lib.vx_event_record(self.ptr, C.c_void_p(stream) if stream else None):
            check("vx_event_record")

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float(0.0)
        if lib.vx_event_elapsed_ms(self.ptr, end.ptr, C.byref(ms)):
            check("vx_event_elapsed_ms")
        return ms.value

    def __del__(self):
        try:
            if self.ptr and self.ptr.value:
                lib.vx_event_destroy(self.ptr)
        except Exception:
            pass


def synchronize_stream(stream: int) -> None:
    if lib.vx_stream_synchronize(C.c_void_p(stream)):
        check("vx_stream_synchronize")


def stwo_twiddles_device(log_n: int):
    """stwo flat twiddle buffers (forward, inverse; 2^(log_n-1) words each) of a
    size-2^log_n domain, built on the device."""
    half = max(1, (1 << log_n) // 2)
    fwd, inv = DeviceBuffer(half), DeviceBuffer(half)
    if lib.vx_stwo_twiddles(log_n, fwd.ptr, inv.ptr, None):
        check("vx_stwo_twiddles")
    synchronize()
    return fwd, inv


class Server:
    """Proof server (Part 5 of the C ABI): jobs of n Fibonacci columns of 2^log_n
    rows, batched per size class on the device under an HBM budget."""

    def __init__(self, budget_bytes: int = 0, n_queries: int = N_QUERIES, log_last_layer: int = LOG_LAST_LAYER,
                 result_bytes: int = 0):
        self.handle = lib.vx_server_create(budget_bytes, result_bytes, n_queries, log_last_layer)
        if not self.handle:
            check("vx_server_create")
            raise HipError("vx_server_create failed")
        self.n_queries = n_queries
        self.log_last_layer = log_last_layer

    def submit(self, log_n: int, ab) -> int:
        flat = (C.c_uint32 * (2 * len(ab)))(*[int(v) % P for pair in ab for v in pair])
        job = lib.vx_server_submit(self.handle, log_n, len(ab), flat)
        if job < 0:
            check("vx_server_submit")
            raise HipError("vx_server_submit failed")
        return job

    def prewarm(self, log_n: int, n_provers: int = 2) -> int:
        """Create provers of a size class before its traffic arrives; returns how many."""
        n = lib.vx_server_prewarm(self.handle, log_n, n_provers)
        if n < 0:
            check("vx_server_prewarm")
            raise HipError("vx_server_prewarm failed")
        return n

    def poll(self) -> int:
        return lib.vx_server_poll(self.handle)

    def set_graphs(self, on: bool) -> None:
        if lib.vx_server_set_graphs(self.handle, 1 if on else 0):
            check("vx_server_set_graphs")

    def set_big_batches(self, max_in_flight: int, log_rows: int = 23) -> None:
        """Run batches of at least 2^log_rows rows at most max_in_flight at a time (0: no limit)."""
        if lib.vx_server_set_big_batches(self.handle, max_in_flight, log_rows):
            check("vx_server_set_big_batches")
            raise HipError("vx_server_set_big_batches failed")

    def set_order(self, smallest_first: bool) -> None:
        """Launch order of size classes in a scheduling step."""
        if lib.vx_server_set_order(self.handle, 1 if smallest_first else 0):
            check("vx_server_set_order")

    def set_stagger(self, on: bool) -> None:
        """Big batches wait on the device for the heavy part of the big batch
        launched before them (the default), or start at once."""
        if lib.vx_server_set_stagger(self.handle, 1 if on else 0):
            check("vx_server_set_stagger")
            raise HipError("vx_server_set_stagger failed")

    def set_copy_kernel(self, on: bool) -> None:
        """Proofs reach the result ring through a copy kernel (True) or the DMA
        engines (False, the default)."""
        if lib.vx_server_set_copy_kernel(self.handle, 1 if on else 0):
            check("vx_server_set_copy_kernel")
            raise HipError("vx_server_set_copy_kernel failed")

    def wait_all(self, raise_on_failure: bool = True) -> bool:
        """Serves until every job has finished; a failed job does not stop the
        others. Returns True when all succeeded; raises HipError when one failed
        unless raise_on_failure is False."""
        ok = lib.vx_server_wait_all(self.handle) == 0
        if not ok and raise_on_failure:
            check("vx_server_wait_all")
            raise HipError("vx_server_wait_all: a job failed")
        return ok

    def inject_fault(self, job: int, col: int = 0) -> None:
        """Test hook: the proof of column col of a queued job is reported failed
        through the device-side failure flag (status -3)."""
        if lib.vx_server_inject_fault(self.handle, job, col):
            check("vx_server_inject_fault")
            raise HipError("vx_server_inject_fault failed")

    def status(self, job: int) -> int:
        return lib.vx_server_job_status(self.handle, job)

    def latency_ms(self, job: int) -> float:
        return lib.vx_server_job_latency_ms(self.handle, job)

    def proof(self, job: int, col: int, log_n: int) -> bytes:
        size = int(lib.vx_krak_proof_size(log_n, self.n_queries, self.log_last_layer))
        buf = (C.c_uint8 * size)()
        if lib.vx_server_proof_copy(self.handle, job, col, buf, size) != size:
            raise HipError(f"job {job}: no proof (status {self.status(job)})")
        return bytes(buf)

    def release(self, job: int) -> None:
        lib.vx_server_release_job(self.handle, job)

    def run_trace(self, arrival_s, log_n, n_cols, ab) -> np.ndarray:
        """Replays arrivals in real time; returns per-job latency in ms."""
        n = len(arrival_s)
        arr_t = np.ascontiguousarray(arrival_s, dtype=np.float64)
        ln = np.ascontiguousarray(log_n, dtype=np.uint32)
        nc = np.ascontiguousarray(n_cols, dtype=np.uint32)
        flat = np.ascontiguousarray(np.asarray(ab, dtype=np.uint64) % P, dtype=np.uint32)
        lat = np.zeros(n, dtype=np.float64)
        if lib.vx_server_run_trace(self.handle, n, arr_t.ctypes.data, ln.ctypes.data, nc.ctypes.data,
                                   flat.ctypes.data, lat.ctypes.data):
            check("vx_server_run_trace")
            raise HipError("vx_server_run_trace failed")
        return lat

    def stats(self) -> dict:
        out = (C.c_uint64 * 6)()
        lib.vx_server_stats(self.handle, out)
        keys = ["batches", "proofs", "jobs_done", "device_bytes", "peak_device_bytes", "budget_bytes"]
        return {k: int(v) for k, v in zip(keys, out)}

    def close(self) -> None:
        if self.handle:
            lib.vx_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ntt_autotune(log_min: int, log_max: int) -> dict:
    """Times every candidate pass split of the NTT for sizes 2^log_min .. 2^log_max
    and installs the fastest; returns {log_n: (ms, [(layers, segment_log), ...])}."""
    best = np.zeros(log_max - log_min + 1, np.float32)
    if lib.vx_ntt_autotune(log_min, log_max, best.ctypes.data):
        check("vx_ntt_autotune")
        raise HipError("vx_ntt_autotune failed")
    return {k: (float(best[k - log_min]), ntt_plan(k)) for k in range(log_min, log_max + 1)}


def ntt_plan(log_n: int) -> list:
    m, c = np.zeros(8, np.uint32), np.zeros(8, np.uint32)
    n = lib.vx_ntt_plan_get(log_n, m.ctypes.data, c.ctypes.data)
    return [(int(m[i]), int(c[i])) for i in range(n)]


def ntt_plan_set(log_n: int, passes) -> None:
    """passes: [(layers, segment_log), ...] with segment_log 0 for the first pass; [] = default."""
    m = np.array([p[0] for p in passes] or [0], np.uint32)
    c = np.array([p[1] for p in passes] or [0], np.uint32)
    if lib.vx_ntt_plan_set(log_n, len(passes), m.ctypes.data, c.ctypes.data):
        check("vx_ntt_plan_set")
        raise HipError("vx_ntt_plan_set failed")


def ntt_lde_plan(log_n: int) -> list:
    """Plan of a blowup-4 evaluation of 2^log_n points (input in the first
    quarter): [(layers, segment_log, replicating), ...]."""
    m, c, r = np.zeros(8, np.uint32), np.zeros(8, np.uint32), np.zeros(8, np.uint32)
    n = lib.vx_ntt_lde_plan_get(log_n, m.ctypes.data, c.ctypes.data, r.ctypes.data)
    return [(int(m[i]), int(c[i]), int(r[i])) for i in range(n)]


def ntt_lde_plan_set(log_n: int, passes) -> None:
    """passes: [(layers, segment_log, replicating), ...]; only the top pass may
    replicate; [] = the general plan of the size."""
    m = np.array([p[0] for p in passes] or [0], np.uint32)
    c = np.array([p[1] for p in passes] or [0], np.uint32)
    r = np.array([p[2] for p in passes] or [0], np.uint32)
    if lib.vx_ntt_lde_plan_set(log_n, len(passes), m.ctypes.data, c.ctypes.data, r.ctypes.data):
        check("vx_ntt_lde_plan_set")
        raise HipError("vx_ntt_lde_plan_set failed")


def merkle_autotune(log_min: int, log_max: int) -> dict:
    """Times whole tree commitments of 2^log_min .. 2^log_max items at every
    stage chunk depth and installs the fastest; returns {log: (ms, depth)}."""
    best = np.zeros(log_max - log_min + 1, np.float32)
    if lib.vx_merkle_autotune(log_min, log_max, best.ctypes.data):
        check("vx_merkle_autotune")
        raise HipError("vx_merkle_autotune failed")
    return {k: (float(best[k - log_min]), int(lib.vx_merkle_stage_plan(k))) for k in range(log_min, log_max + 1)}


def merkle_stage_plan(log_items: int) -> int:
    """Tuned stage chunk depth for 2^log_items items per launch (-1: default rule)."""
    return int(lib.vx_merkle_stage_plan(log_items))


def merkle_stage_plan_set(log_items: int, logc: int) -> None:
    if lib.vx_merkle_stage_plan_set(log_items, logc):
        check("vx_merkle_stage_plan_set")
        raise HipError("vx_merkle_stage_plan_set failed")


class DeepQuotient:
    """Fused stwo DEEP quotient (vx_deep_*): QuotientOps::accumulate_numerators and
    compute_quotients_and_combine (quotient_ops.rs:35-223) in one launch."""

    def __init__(self, max_cols: int, max_points: int = 2):
        self.handle = lib.vx_deep_create(max_cols, max_points)
        if not self.handle:
            check("vx_deep_create")
            raise HipError("vx_deep_create failed")

    def quotient(self, col_addrs, col_logs, points, line_coeffs, lifting_log: int, domain_xs: int, domain_ys: int,
                 outs, stream: int = 0) -> None:
        """col_addrs: device addresses; points: (n_points, 8) words; line_coeffs:
        (n_points, n_cols, 12) words (a, b, c); outs: four device addresses."""
        n = len(col_addrs)
        ptrs = (C.c_void_p * n)(*col_addrs)
        logs = np.ascontiguousarray(col_logs, dtype=np.uint32)
        pts = np.ascontiguousarray(points, dtype=np.uint32)
        lc = np.ascontiguousarray(line_coeffs, dtype=np.uint32)
        if lib.vx_deep_quotient(self.handle, ptrs, logs.ctypes.data, n, pts.size // 8, pts.ctypes.data,
                                lc.ctypes.data, lifting_log, C.c_void_p(domain_xs), C.c_void_p(domain_ys),
                                *[C.c_void_p(o) for o in outs], C.c_void_p(stream) if stream else None):
            check("vx_deep_quotient")
            raise HipError("vx_deep_quotient failed")

    def close(self) -> None:
        if self.handle:
            lib.vx_deep_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def circle_domain_points(log_size: int, stream: int = 0):
    """(xs, ys) device buffers: the canonic circle domain at bit-reversed rows."""
    xs, ys = DeviceBuffer(1 << log_size), DeviceBuffer(1 << log_size)
    if lib.vx_circle_domain_points(log_size, xs.ptr, ys.ptr, C.c_void_p(stream) if stream else None):
        check("vx_circle_domain_points")
        raise HipError("vx_circle_domain_points failed")
    return xs, ys


NONCE_NONE = (1 << 64) - 1


def grind(state, pow_bits: int, m31_output: bool = False, start: int = 0, count: int = NONCE_NONE):
    """Proof-of-work nonce for a Blake2s channel state (8 words, host), as
    GrindOps::grind (grind_ops.rs:125-180): the smallest nonce of
    [start, start + count) whose Blake2s(prefixed_digest || nonce) has >= pow_bits
    trailing zeros, searched on the device in one launch; None if the range has none."""
    ch = DeviceBuffer.from_host(np.concatenate([np.asarray(state, np.uint32), np.zeros(8, np.uint32)]))
    res = DeviceBuffer(2)
    if lib.vx_grind_blake2s(ch.ptr, pow_bits, int(m31_output), start, count, res.ptr, 0, None):
        check("vx_grind_blake2s")
        raise HipError("vx_grind_blake2s failed")
    synchronize()
    lo, hi = res.to_host(2)
    nonce = int(lo) | (int(hi) << 32)
    return None if nonce == NONCE_NONE else nonce


def grind_device(channel_addr: int, pow_bits: int, result_addr: int, m31_output: bool = False, start: int = 0,
                 count: int = NONCE_NONE, mix: bool = False, stream: int = 0) -> None:
    """Asynchronous device PoW step on a device transcript (first 8 words = state):
    the nonce goes to result_addr (u64, 2^64 - 1 if none); mix = True then mixes it
    into the transcript as Channel::mix_u64 (channel.rs:117-128)."""
    if lib.vx_grind_blake2s(C.c_void_p(channel_addr), pow_bits, int(m31_output), start, count,
                            C.c_void_p(result_addr), int(mix), C.c_void_p(stream) if stream else None):
        check("vx_grind_blake2s")
        raise HipError("vx_grind_blake2s failed")


def measure_valu_peak() -> float:
    """int32 VALU issue rate of the device, T lane-ops/s (microbenchmark)."""
    tops = C.c_double(0.0)
    if lib.vx_measure_valu_peak(C.byref(tops), None):
        check("vx_measure_valu_peak")
        raise HipError("vx_measure_valu_peak failed")
    return tops.value


def measure_hash_latency() -> tuple:
    """(quad-cooperative, one-lane) latency of one Merkle node hash in a chain
    of dependent nodes on one wave, in ns (microbenchmark)."""
    q, l = C.c_double(0.0), C.c_double(0.0)
    if lib.vx_measure_hash_latency(C.byref(q), C.byref(l)):
        check("vx_measure_hash_latency")
        raise HipError("vx_measure_hash_latency failed")
    return q.value, l.value
