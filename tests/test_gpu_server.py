"""Proof server: jobs of mixed sizes, batched per size class under an HBM
budget, give every proof bit-identical to proving it alone (and to the
oracle). Runs on the MI355X."""

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu
P = (1 << 31) - 1


def jobs_spec(seed, n, sizes, max_cols):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        log_n = int(rng.choice(sizes))
        cols = int(rng.integers(1, max_cols + 1))
        out.append((log_n, [(int(rng.integers(0, P)), int(rng.integers(0, P))) for _ in range(cols)]))
    return out


def test_server_mixed_jobs_match_oracle(vx):
    spec = jobs_spec(1, 12, [4, 6, 9, 11], 3)
    srv = vx.Server(0, 64, 3)
    ids = [srv.submit(log_n, ab) for log_n, ab in spec]
    srv.wait_all()
    solo = {}
    for jid, (log_n, ab) in zip(ids, spec):
        assert srv.status(jid) == 1
        assert srv.latency_ms(jid) > 0
        for c, (a, b) in enumerate(ab):
            exp = O.prove(a, b, log_n, 64, 3)
            assert srv.proof(jid, c, log_n) == exp, (jid, c)
            if log_n not in solo:
                solo[log_n] = vx.Prover(log_n, 64, 3)
            assert solo[log_n].prove(a, b) == exp
    st = srv.stats()
    assert st["proofs"] == sum(len(ab) for _, ab in spec)
    assert st["batches"] < st["proofs"]  # proofs of a class share launches
    for p in solo.values():
        p.close()
    srv.close()


def test_server_trace_replay(vx):
    spec = jobs_spec(2, 16, [5, 8, 10], 4)
    arrivals = np.sort(np.random.default_rng(3).uniform(0, 0.05, len(spec)))
    srv = vx.Server(0, 80, 3)
    lat = srv.run_trace(arrivals, [s[0] for s in spec], [len(s[1]) for s in spec],
                        [v for s in spec for pair in s[1] for v in pair])
    assert lat.shape == (len(spec),) and np.all(lat > 0)
    for jid, (log_n, ab) in enumerate(spec):
        for c, (a, b) in enumerate(ab):
            assert srv.proof(jid, c, log_n) == O.prove(a, b, log_n, 80, 3), (jid, c)
        srv.release(jid)
    srv.close()


def test_server_budget_too_small_fails_cleanly(vx):
    srv = vx.Server(1 << 20, 80, 3)  # 1 MiB cannot hold a 2^12 prover
    jid = srv.submit(12, [(1, 2)])
    with pytest.raises(vx.HipError):
        srv.wait_all()
    assert srv.status(jid) == -4
    srv.close()


def test_server_small_result_ring_with_releases(vx):
    # room for three proofs: batches wait for finished jobs to be released
    size = int(vx.lib.vx_krak_proof_size(6, 80, 3))
    srv = vx.Server(0, 80, 3, result_bytes=3 * size)
    spec = [(6, [(i + 1, 2 * i + 5)]) for i in range(10)]
    ids = [srv.submit(l, ab) for l, ab in spec]
    checked = set()
    for _ in range(200000):
        srv.poll()
        for jid in ids:
            if jid not in checked and srv.status(jid) == 1:
                (a, b), = spec[jid][1]
                assert srv.proof(jid, 0, 6) == O.prove(a, b, 6, 80, 3)
                srv.release(jid)
                checked.add(jid)
        if len(checked) == len(ids):
            break
    assert len(checked) == len(ids)
    srv.close()


def test_server_ring_full_without_release_fails_cleanly(vx):
    size = int(vx.lib.vx_krak_proof_size(6, 80, 3))
    srv = vx.Server(0, 80, 3, result_bytes=2 * size)
    ids = [srv.submit(6, [(1, 1)]) for _ in range(5)]
    with pytest.raises(vx.HipError):
        srv.wait_all()
    st = [srv.status(i) for i in ids]
    assert st.count(1) == 2 and all(x == -5 for x in st if x != 1)
    srv.close()


def test_prewarm_makes_provers_ahead_of_traffic(vx):
    """vx_server_prewarm creates up to two provers of a class within the budget;
    proofs served afterwards are the oracle's."""
    import oracle_lib as O

    srv = vx.Server(0, 80, 3)
    assert srv.prewarm(10, 2) == 2
    assert srv.prewarm(10, 5) == 0  # already at the per-class limit
    held = srv.stats()["device_bytes"]
    assert held > 0
    ab = [(3, 5), (7, 11), (13, 17)]
    job = srv.submit(10, ab)
    srv.wait_all()
    for c, (a, b) in enumerate(ab):
        assert srv.proof(job, c, 10) == O.prove(a, b, 10, 80, 3)
    assert srv.stats()["batches"] >= 1
    srv.close()
    # a budget that holds no prover makes none
    small = vx.Server(1 << 20, 80, 3)
    assert small.prewarm(12, 2) == 0
    small.close()


def test_injected_fault_fails_one_job_alone(vx):
    """A proof reported failed by the device (the injected fault raises the
    transcript failure flag of one proof inside a shared batch) fails its job
    alone: every other job, including the proofs batched with it, completes
    with the oracle's bytes."""
    spec = jobs_spec(11, 10, [5, 7], 3)
    srv = vx.Server(0, 80, 3)
    ids = [srv.submit(log_n, ab) for log_n, ab in spec]
    bad = ids[3]
    srv.inject_fault(bad, len(spec[3][1]) - 1)
    assert srv.wait_all(raise_on_failure=False) is False
    for jid, (log_n, ab) in zip(ids, spec):
        if jid == bad:
            assert srv.status(jid) == -3
            with pytest.raises(vx.HipError):
                srv.proof(jid, 0, log_n)
            continue
        assert srv.status(jid) == 1, jid
        for c, (a, b) in enumerate(ab):
            assert srv.proof(jid, c, log_n) == O.prove(a, b, log_n, 80, 3), (jid, c)
    # the failed proof's arena serves the next batch normally
    again = srv.submit(spec[3][0], spec[3][1])
    assert srv.wait_all()
    for c, (a, b) in enumerate(spec[3][1]):
        assert srv.proof(again, c, spec[3][0]) == O.prove(a, b, spec[3][0], 80, 3)
    srv.close()


def test_class_over_budget_fails_alone(vx):
    """Admission failure of one size class (its prover cannot fit the budget)
    fails that class's job with -4 while the jobs of a class that fits finish
    with the oracle's bytes."""
    small = 6
    # provers hold 2^(24 - log_n) proofs, between 2 and 1024
    fit = int(vx.lib.vx_prover_estimate_bytes(small, 80, 3, 1024))
    budget = 2 * fit + (1 << 20)
    assert fit > 0 and int(vx.lib.vx_prover_estimate_bytes(16, 80, 3, 256)) > budget
    srv = vx.Server(budget, 80, 3)
    ok_jobs = [srv.submit(small, [(i + 3, 7 * i + 1)]) for i in range(4)]
    big = srv.submit(16, [(5, 9)])
    assert srv.wait_all(raise_on_failure=False) is False
    assert srv.status(big) == -4
    for i, jid in enumerate(ok_jobs):
        assert srv.status(jid) == 1
        assert srv.proof(jid, 0, small) == O.prove(i + 3, 7 * i + 1, small, 80, 3)
    srv.close()


def test_inject_fault_argument_errors(vx):
    srv = vx.Server(0, 80, 3)
    with pytest.raises(vx.HipError):
        srv.inject_fault(0, 0)  # no such job
    jid = srv.submit(5, [(1, 2)])
    with pytest.raises(vx.HipError):
        srv.inject_fault(jid, 1)  # no such column
    srv.wait_all()
    with pytest.raises(vx.HipError):
        srv.inject_fault(jid, 0)  # already finished
    srv.close()


def test_unsupported_protocol_is_refused_at_creation(vx):
    """More than 4096 queries or a last layer above 2^10 can never make a
    prover: the server refuses them at creation instead of spinning on jobs it
    cannot serve; an unsupported size class is refused at submission."""
    with pytest.raises(vx.HipError):
        vx.Server(0, 5000, 3)
    with pytest.raises(vx.HipError):
        vx.Server(0, 80, 11)
    srv = vx.Server(0, 80, 3)
    with pytest.raises(vx.HipError):
        srv.submit(30, [(1, 2)])  # 2^32 points: beyond the circle group
    with pytest.raises(vx.HipError):
        srv.submit(1, [(1, 2)])
    srv.close()


@pytest.mark.parametrize("smallest_first", [True, False])
def test_big_batch_cap_and_order_keep_bytes(vx, smallest_first):
    """With at most one batch of >= 2^12 rows in flight and either class
    order, every proof still equals the oracle's and every job completes."""
    spec = jobs_spec(21, 12, [4, 6, 8], 4)
    srv = vx.Server(0, 80, 3)
    srv.set_order(smallest_first)
    srv.set_big_batches(1, 12)
    ids = [srv.submit(log_n, ab) for log_n, ab in spec]
    assert srv.wait_all()
    for jid, (log_n, ab) in zip(ids, spec):
        assert srv.status(jid) == 1
        for c, (a, b) in enumerate(ab):
            assert srv.proof(jid, c, log_n) == O.prove(a, b, log_n, 80, 3), (jid, c)
    with pytest.raises(vx.HipError):
        srv.set_big_batches(1, 41)
    srv.close()


@pytest.mark.parametrize("stagger", [True, False])
def test_staggered_big_batches_keep_bytes(vx, stagger):
    """Big batches (here >= 2^10 rows) wait for the heavy part of the big batch
    launched before them, or start at once: proofs equal the oracle's either
    way, over two passes (the second replays captured graphs)."""
    spec = jobs_spec(22, 10, [5, 7, 9], 3)
    srv = vx.Server(0, 80, 3)
    srv.set_graphs(True)
    srv.set_big_batches(0, 10)
    srv.set_stagger(stagger)
    for _ in range(2):
        ids = [srv.submit(log_n, ab) for log_n, ab in spec]
        assert srv.wait_all()
        for jid, (log_n, ab) in zip(ids, spec):
            assert srv.status(jid) == 1
            for c, (a, b) in enumerate(ab):
                assert srv.proof(jid, c, log_n) == O.prove(a, b, log_n, 80, 3), (jid, c)
            srv.release(jid)
    srv.close()
