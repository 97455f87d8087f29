"""Batched proving: many proofs per launch sequence (one arena per proof) and
HIP-graph replay give the same KRAK bytes as the oracle and as one proof at a
time. Runs on the MI355X."""

import hashlib

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu
P = (1 << 31) - 1


def pairs(seed, n):
    rng = np.random.default_rng(seed)
    return [(int(rng.integers(0, P)), int(rng.integers(0, P))) for _ in range(n)]


@pytest.mark.parametrize("log_n,count,cap", [(4, 5, 8), (8, 7, 7), (12, 3, 4), (15, 2, 2)])
def test_batch_matches_oracle_and_solo(vx, log_n, count, cap):
    ab = pairs(log_n * 100 + count, count)
    bp = vx.Prover(log_n, 80, 3, capacity=cap)
    got = bp.prove_batch(ab)
    solo = vx.Prover(log_n, 80, 3)
    for i, (a, b) in enumerate(ab):
        exp = O.prove(a, b, log_n, 80, 3)
        assert got[i] == exp, (log_n, i)
        assert solo.prove(a, b) == exp
    bp.close()
    solo.close()


def test_batch_of_device_traces_partial(vx):
    log_n, cap = 10, 6
    ab = pairs(5, 4)
    traces = [vx.fibonacci_trace_device(a, b, log_n) for a, b in ab]
    bp = vx.Prover(log_n, 64, 4, capacity=cap)
    bp.prove_batch_device([t.addr for t in traces], ab)
    for i, (a, b) in enumerate(ab):
        assert bp.proof_bytes(i) == O.prove(a, b, log_n, 64, 4)
    # a smaller batch next: slots 0..1 are rewritten, the rest keep the earlier proofs
    bp.prove_batch_device([traces[3].addr, traces[2].addr], [ab[3], ab[2]])
    assert bp.proof_bytes(0) == O.prove(*ab[3], log_n, 64, 4)
    assert bp.proof_bytes(1) == O.prove(*ab[2], log_n, 64, 4)
    assert bp.proof_bytes(3) == O.prove(*ab[3], log_n, 64, 4)
    bp.close()


def test_graph_replays_match_eager(vx):
    log_n, cap = 9, 5
    bp = vx.Prover(log_n, 80, 3, capacity=cap)
    bp.set_graph(True)
    scratch = vx.DeviceBuffer(1 << 12)
    for rep in range(4):
        ab = pairs(40 + rep, cap if rep % 2 == 0 else 3)  # new inputs; batch count changes -> new capture
        got = bp.prove_batch(ab)
        for i, (a, b) in enumerate(ab):
            assert got[i] == O.prove(a, b, log_n, 80, 3), (rep, i)
        # eager work of the program between replays
        vx.lib.vx_memset(scratch.ptr, rep, 4 << 12)
        vx.synchronize()
    # the same inputs replayed twice in a row
    ab = pairs(99, cap)
    h1 = [hashlib.sha256(x).hexdigest() for x in bp.prove_batch(ab)]
    h2 = [hashlib.sha256(x).hexdigest() for x in bp.prove_batch(ab)]
    assert h1 == h2 == [hashlib.sha256(O.prove(a, b, log_n, 80, 3)).hexdigest() for a, b in ab]
    bp.set_graph(False)
    assert bp.prove_batch(ab[:2]) == [O.prove(a, b, log_n, 80, 3) for a, b in ab[:2]]
    bp.close()


@pytest.mark.parametrize("lde_plan", [[(11, 0, 1)], [(6, 0, 0), (5, 5, 1)]])
def test_graph_replays_with_lde_bridge(vx, lde_plan):
    """the bridged LDE (one contiguous bridge pass, or a strided one above a
    contiguous pass) captured in a graph and replayed with eager work between
    replays; every replay gives the oracle's proofs"""
    log_n, cap = 9, 4
    vx.ntt_lde_plan_set(log_n + 2, lde_plan)
    try:
        bp = vx.Prover(log_n, 80, 3, capacity=cap)
        bp.set_graph(True)
        scratch = vx.DeviceBuffer(1 << 12)
        ab = pairs(7, cap)
        want = [O.prove(a, b, log_n, 80, 3) for a, b in ab]
        for rep in range(3):
            assert bp.prove_batch(ab) == want, rep
            vx.lib.vx_memset(scratch.ptr, rep, 4 << 12)
            vx.synchronize()
        bp.close()
    finally:
        vx.lib.vx_ntt_plan_reset()


@pytest.mark.gpu
@pytest.mark.parametrize("log_n,nq,last,cap", [(10, 80, 3, 4), (14, 64, 4, 2), (18, 100, 4, 1), (20, 80, 3, 3)])
def test_device_bytes_equal_estimate(vx, log_n, nq, last, cap):
    p = vx.Prover(log_n, nq, last, capacity=cap)
    try:
        assert p.device_bytes == vx.lib.vx_prover_estimate_bytes(log_n, nq, last, cap)
    finally:
        p.close()


@pytest.mark.gpu
@pytest.mark.parametrize("copy_kernel", [False, True])
def test_enqueue_copy_pinned_and_pageable(vx, copy_kernel):
    """vx_prover_enqueue_copy writes a batch's proofs and failure flags into
    pinned memory (with the DMA engines, or with a device kernel once
    vx_prover_set_copy_kernel chose it) and into other host memory with a DMA
    copy; every way gives the bytes of proof_bytes."""
    import ctypes as C

    ab = pairs(11, 5)
    p = vx.Prover(9, 80, 3, capacity=8)
    try:
        assert vx.lib.vx_prover_set_copy_kernel(p.handle, 1 if copy_kernel else 0) == 0
        p.prove_batch_async(ab)
        vx.synchronize_stream(p.stream)
        want = [p.proof_bytes(i) for i in range(len(ab))]
        size = p.proof_size * len(ab)
        # pinned
        hp, hf = C.c_void_p(), C.c_void_p()
        assert vx.lib.vx_malloc_host(C.byref(hp), size) == 0 and vx.lib.vx_malloc_host(C.byref(hf), 4 * len(ab)) == 0
        try:
            C.memset(hp, 0xAB, size)
            C.memset(hf, 0xFF, 4 * len(ab))
            assert vx.lib.vx_prover_enqueue_copy(p.handle, len(ab), hp, hf, None) == 0
            vx.synchronize_stream(p.stream)
            got = C.string_at(hp, size)
            flags = np.frombuffer(C.string_at(hf, 4 * len(ab)), dtype=np.uint32)
        finally:
            vx.lib.vx_free_host(hp)
            vx.lib.vx_free_host(hf)
        assert [got[i * p.proof_size:(i + 1) * p.proof_size] for i in range(len(ab))] == want
        assert (flags == 0).all()
        # pageable: DMA copy
        buf = (C.c_uint8 * size)()
        fl = (C.c_uint32 * len(ab))(*([7] * len(ab)))
        assert vx.lib.vx_prover_enqueue_copy(p.handle, len(ab), buf, fl, None) == 0
        vx.synchronize_stream(p.stream)
        got2 = bytes(buf)
        assert [got2[i * p.proof_size:(i + 1) * p.proof_size] for i in range(len(ab))] == want
        assert list(fl) == [0] * len(ab)
    finally:
        p.close()
