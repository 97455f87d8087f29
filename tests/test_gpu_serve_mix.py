"""The served stream's own workload shape: Poisson arrivals of Fibonacci jobs
with 1-34 columns and log_n drawn from 12-18, with jobs of 2^20-2^22 rows mixed
in, replayed in real time through the server (continuous batching under the
HBM budget). Each served proof must be byte-identical to the same proof made
alone: by a standalone capacity-1 prover for every column, by the oracle for
the first and last column of every job up to 2^18 rows, and the restated
verifier must accept every proof of the larger jobs. Runs on the MI355X."""

import numpy as np
import pytest

import inputs
import oracle_lib as O

pytestmark = pytest.mark.gpu
P = (1 << 31) - 1


def served_jobs(seed: int, n_small: int):
    """n_small jobs with log_n uniform in [12, 18] (log-uniform sizes) and 1-34
    columns, plus one job each of 2^20, 2^21 and 2^22 rows with 1-3 columns,
    placed among them; (a, b) per column from splitmix64."""
    r = inputs.splitmix64(seed, 2 * n_small + 3)
    jobs = []
    for j in range(n_small):
        jobs.append((12 + int(r[2 * j] % np.uint64(7)), 1 + int(r[2 * j + 1] % np.uint64(34))))
    for k, log_n in enumerate((20, 21, 22)):
        jobs.insert(2 + 4 * k, (log_n, 1 + int(r[2 * n_small + k] % np.uint64(3))))
    v = inputs.splitmix64(seed + 1, 2 * sum(c for _, c in jobs)) % np.uint64(P)
    out, pos = [], 0
    for log_n, cols in jobs:
        out.append((log_n, [(int(v[pos + 2 * c]), int(v[pos + 2 * c + 1])) for c in range(cols)]))
        pos += 2 * cols
    return out


@pytest.mark.parametrize("rate", [100.0, 1e6])
def test_served_stream_matches_alone(vx, rate):
    spec = served_jobs(51, 10)
    assert {l for l, _ in spec} >= {20, 21, 22} and max(len(ab) for _, ab in spec) > 8
    gaps = -np.log1p(-(inputs.splitmix64(52, len(spec)) >> np.uint64(11)).astype(np.float64) / float(1 << 53))
    arrivals = np.cumsum(gaps / rate)
    arrivals -= arrivals[0]
    print(f"served mix: {len(spec)} jobs, {sum(len(ab) for _, ab in spec)} proofs, rate {rate} jobs/s", flush=True)
    srv = vx.Server(0, 80, 3)
    lat = srv.run_trace(arrivals, [l for l, _ in spec], [len(ab) for _, ab in spec],
                        [v for _, ab in spec for pair in ab for v in pair])
    assert lat.shape == (len(spec),) and np.all(lat > 0)
    st = srv.stats()
    assert st["batches"] < st["proofs"]  # proofs shared launches
    solo = {}
    for jid, (log_n, ab) in enumerate(spec):
        assert srv.status(jid) == 1, jid
        if log_n not in solo:
            solo[log_n] = vx.Prover(log_n, 80, 3)
        for c, (a, b) in enumerate(ab):
            got = srv.proof(jid, c, log_n)
            assert got == solo[log_n].prove(a, b), (jid, c, log_n)
            if log_n <= 18 and c in (0, len(ab) - 1):
                assert got == O.prove(a, b, log_n, 80, 3), (jid, c, log_n)
            if log_n >= 20:
                assert O.verify(got, 80, 3) == 0, (jid, c, log_n)
        srv.release(jid)
    for p in solo.values():
        p.close()
    srv.close()
