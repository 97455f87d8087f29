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
