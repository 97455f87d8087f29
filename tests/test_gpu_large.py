"""The largest proofs on the MI355X: 2^26 and 2^28 rows, and 2^29 rows, whose
2^31-point evaluation domain is the whole circle group subgroup(31)
(prover.rs:626-630, 754-757; verifier.rs:160-165). Each proof's commitments
(trace and quotient roots, every FRI root, the last layer and the query
indices) must equal the oracle's memory-lean summary committed in
tests/golden/large.json, and the restated verifier must accept the proof.
Equal roots bind every decommitted value and path, so together these pin the
proof bytes."""

import json
import os

import pytest

import krak
import oracle_lib as O

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "large.json")


def golden(log_n):
    with open(GOLDEN) as f:
        return next(c for c in json.load(f) if c["log_n"] == log_n)


@pytest.mark.parametrize("log_n", [26, 28, 29])
def test_largest_proofs_match_oracle_summary(vx, log_n):
    c = golden(log_n)
    print(f"large proof: log_n={log_n}, a={c['a']}, b={c['b']}", flush=True)
    p = vx.Prover(log_n, c["n_queries"], c["log_last"])
    try:
        got = p.prove(c["a"], c["b"])
    finally:
        p.close()
    pf = krak.parse(got)
    assert pf.log_n == log_n
    assert pf.trace_root == c["trace_root"]
    assert pf.quotient_root == c["quotient_root"]
    assert pf.fri_roots == c["fri_roots"]
    assert pf.last_layer == c["last_layer"]
    assert pf.queries == c["queries"]
    assert O.verify(got, c["n_queries"], c["log_last"]) == 0
