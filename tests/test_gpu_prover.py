"""Parity of the fused device-resident prover (include/vortex_hip.h, Part 3):
identical KRAK v1 bytes to the oracle, accepted by the restated verifier, on
the reference verifier grid, the plain CPU workload and the plain GPU workload
(2^24 rows, against the committed golden summary)."""

import hashlib
import json
import os

import numpy as np
import pytest

import krak
import oracle_lib as O

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "proofs.json")


def golden(label):
    with open(GOLDEN) as f:
        return next(c for c in json.load(f) if c["label"] == label)


@pytest.mark.parametrize("log_n", [4, 5, 6, 8])
@pytest.mark.parametrize("ab", [(1, 1), (2, 3), (42, 99), (1000, 2000)])
def test_proof_bytes_grid(vx, log_n, ab):
    got = vx.prove(ab[0], ab[1], log_n)
    assert got == O.prove(ab[0], ab[1], log_n)
    assert O.verify(got) == 0


@pytest.mark.parametrize("log_n,nq,last", [(10, 80, 3), (12, 64, 4), (13, 100, 4), (14, 80, 3)])
def test_proof_bytes_medium(vx, log_n, nq, last):
    got = vx.prove(5, 8, log_n, nq, last)
    assert got == O.prove(5, 8, log_n, nq, last)
    assert O.verify(got, nq, last) == 0


@pytest.mark.parametrize("label", ["cpu_plain", "cpu_plain_reference_protocol"])
def test_plain_cpu_workload(vx, label):
    c = golden(label)
    got = vx.prove(c["a"], c["b"], c["log_n"], c["n_queries"], c["log_last"])
    assert hashlib.sha256(got).hexdigest() == c["sha256"]
    assert O.verify(got, c["n_queries"], c["log_last"]) == 0


def test_prover_reuse_and_lde(vx):
    p = vx.Prover(11, 80, 3)
    first = p.prove(3, 4)
    other = p.prove(9, 9)
    again = p.prove(3, 4)
    assert first == again and first != other
    ev = p.eval_words()
    assert np.array_equal(ev, O.lde(O.fibonacci_trace(3, 4, 11), 11))
    # first FRI layer equals the circle fold of the quotient (fri.rs:163-191)
    pf = krak.parse(first)
    assert pf.trace_root == O.merkle_root([ev]).tolist()
    p.close()


def test_device_trace_generator(vx):
    for log_n in (4, 10, 17):
        d = vx.fibonacci_trace_device(12345, 678910, log_n)
        assert np.array_equal(d.to_host(), O.fibonacci_trace(12345, 678910, log_n))


def test_plain_gpu_workload_log20(vx):
    got = vx.prove(77, 88, 20, 100, 4)
    exp = O.prove(77, 88, 20, 100, 4)
    assert got == exp


def test_plain_gpu_workload_log24(vx):
    c = golden("gpu_plain")
    p = vx.Prover(24, c["n_queries"], c["log_last"])
    got = p.prove(c["a"], c["b"])
    pf = krak.parse(got)
    assert pf.trace_root == c["trace_root"]
    assert pf.quotient_root == c["quotient_root"]
    assert pf.fri_roots == c["fri_roots"]
    assert pf.last_layer == c["last_layer"]
    assert pf.queries == c["queries"]
    assert hashlib.sha256(got).hexdigest() == c["sha256"]
    ev = p.eval_words()
    assert ev[:64].tolist() == c["lde_head"] and ev[-64:].tolist() == c["lde_tail"]
    assert hashlib.sha256(ev.tobytes()).hexdigest() == c["lde_sha256"]
    assert O.verify(got, c["n_queries"], c["log_last"]) == 0
    p.close()


def test_two_provers_on_two_streams_concurrently(vx):
    """Proofs enqueued alternately on two provers (two streams, running
    concurrently on the device) are the oracle's, each prover's last proof
    identical."""
    log_n = 13
    want = O.prove(7, 11, log_n, 80, 3)
    p1, p2 = vx.Prover(log_n, 80, 3), vx.Prover(log_n, 80, 3)
    t1 = vx.fibonacci_trace_device(7, 11, log_n)
    t2 = vx.fibonacci_trace_device(7, 11, log_n)
    try:
        for i in range(6):
            (p1, p2)[i % 2].prove_device((t1, t2)[i % 2].addr, 7, 11)
        vx.synchronize()
        assert p1.proof_bytes() == want and p2.proof_bytes() == want
    finally:
        p1.close()
        p2.close()
        t1.free()
        t2.free()


def test_provers_used_from_two_threads(vx):
    """One prover per thread (the library keeps its error state and launch
    stream per thread): concurrent proofs from two host threads are the
    oracle's."""
    import threading

    log_n = 12
    cases = [(3, 5), (8, 13)]
    want = [O.prove(a, b, log_n, 80, 3) for a, b in cases]
    got = [None, None]

    def work(k):
        p = vx.Prover(log_n, 80, 3)
        try:
            a, b = cases[k]
            for _ in range(4):
                got[k] = p.prove(a, b)
        finally:
            p.close()

    th = [threading.Thread(target=work, args=(k,)) for k in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=60)
    assert got == want
