"""Proofs of traces in host memory (vx_prover_prove_batch_host): pageable and
pinned traces, uploads of one call overlapping the proving of the previous
one, mixed with device-trace calls and graph replays, give the oracle's KRAK
bytes. Runs on the MI355X."""

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu
P = (1 << 31) - 1


def pairs(seed, n):
    rng = np.random.default_rng(seed)
    return [(int(rng.integers(0, P)), int(rng.integers(0, P))) for _ in range(n)]


def run_calls(vx, bp, calls, log_n, pinned):
    """Enqueue every call back to back (no host wait between them), copying each
    call's proofs to its own pinned host buffer on the prover's stream."""
    size = bp.proof_size
    outs, keep = [], []
    for ab in calls:
        traces = []
        for a, b in ab:
            t = O.fibonacci_trace(a, b, log_n)
            if pinned:
                pb = vx.PinnedBuffer(t.size)
                pb.array[:] = t
                keep.append(pb)
                traces.append(pb.addr)
            else:
                traces.append(t)
        bp.prove_batch_host(traces, ab)
        out = vx.PinnedBuffer(size * len(ab) // 4)
        if vx.lib.vx_prover_enqueue_copy(bp.handle, len(ab), out.ptr, None, None):
            vx.check("vx_prover_enqueue_copy")
        outs.append(out)
    vx.synchronize_stream(bp.stream)
    vx.check("host traces")
    return [[out.array.tobytes()[i * size:(i + 1) * size] for i in range(len(ab))] for out, ab in zip(outs, calls)]


@pytest.mark.parametrize("pinned", [False, True])
@pytest.mark.parametrize("graph", [False, True])
def test_host_trace_calls_back_to_back(vx, pinned, graph):
    log_n, cap = 11, 3
    calls = [pairs(1, 3), pairs(2, 1), pairs(3, 2), pairs(4, 3), pairs(5, 1)]
    bp = vx.Prover(log_n, 80, 3, capacity=cap)
    bp.set_graph(graph)
    got = run_calls(vx, bp, calls, log_n, pinned)
    for c, ab in enumerate(calls):
        for i, (a, b) in enumerate(ab):
            assert got[c][i] == O.prove(a, b, log_n, 80, 3), (c, i)
    bp.close()


def test_host_and_device_calls_interleaved(vx):
    log_n = 13
    bp = vx.Prover(log_n, 64, 4, capacity=2)
    ab = pairs(7, 6)
    t_host = [O.fibonacci_trace(a, b, log_n) for a, b in ab]
    d0 = vx.fibonacci_trace_device(*ab[2], log_n)
    bp.prove_batch_host(t_host[:2], ab[:2])
    first = [bp.proof_bytes(0), bp.proof_bytes(1)]
    bp.prove_batch_device([d0.addr], [ab[2]])
    mid = bp.proof_bytes(0)
    bp.prove_batch_host(t_host[3:5], ab[3:5])
    bp.prove_batch_host(t_host[5:6], ab[5:6])  # the other slot, then the first again
    bp.prove_batch_host(t_host[:1], ab[:1])
    last = bp.proof_bytes(0)
    assert first == [O.prove(a, b, log_n, 64, 4) for a, b in ab[:2]]
    assert mid == O.prove(*ab[2], log_n, 64, 4)
    assert last == O.prove(*ab[0], log_n, 64, 4)
    bp.close()


def test_host_trace_full_size_matches_device_trace(vx):
    """2^20 rows: the proof of a pinned host trace equals the proof of the same
    trace generated on the device (a size-independent property: same input,
    same bytes), and the oracle's proof."""
    log_n = 20
    a, b = 12345, 67890
    bp = vx.Prover(log_n, 80, 3)
    t = vx.PinnedBuffer(1 << log_n)
    t.array[:] = O.fibonacci_trace(a, b, log_n)
    bp.prove_batch_host([t.addr], [(a, b)])
    from_host = bp.proof_bytes(0)
    d = vx.fibonacci_trace_device(a, b, log_n)
    bp.prove_device(d.addr, a, b)
    assert from_host == bp.proof_bytes(0) == O.prove(a, b, log_n, 80, 3)
    bp.close()
