"""The oracle's whole proof: bytes against the independent Python model, the
reference verifier's accept/reject behaviour, and the committed golden fixtures."""

import hashlib
import json
import os

import pytest

import krak
import pyref

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("a,b,log_n,nq,last", [(1, 1, 4, 80, 3), (2, 3, 4, 80, 3), (42, 99, 5, 80, 3),
                                               (7, 11, 4, 64, 4), (1000, 2000, 5, 100, 4)])
def test_oracle_proof_equals_python_model(oracle, a, b, log_n, nq, last):
    assert oracle.prove(a, b, log_n, nq, last) == pyref.prove(a, b, log_n, nq, last)


@pytest.mark.parametrize("log_n", [4, 5, 6, 8])
@pytest.mark.parametrize("ab", [(1, 1), (2, 3), (42, 99), (1000, 2000)])
def test_verifier_accepts(oracle, log_n, ab):
    # verifier.rs:369-384 grid
    proof = oracle.prove(ab[0], ab[1], log_n)
    assert oracle.verify(proof) == 0
    pf = krak.parse(proof)
    assert len(pf.fri_roots) == log_n + 2 - 3
    assert len(pf.last_layer) == 8
    assert krak.serialize(pf) == proof


def _tampered(proof: bytes, fn) -> bytes:
    pf = krak.parse(proof)
    fn(pf)
    return krak.serialize(pf)


@pytest.mark.parametrize("what", ["trace_value", "quotient_value", "fri_value", "commitment", "fri_layers",
                                  "zero_commitment", "last_layer", "query", "path", "truncated"])
def test_verifier_rejects_tampering(oracle, what):
    # verifier.rs:324-366 plus extra cases
    proof = oracle.prove(1, 1, 6)
    if what == "truncated":
        bad = proof[:-4]
    else:
        def mutate(pf):
            if what == "trace_value":
                pf.trace.values[0] ^= 1
            elif what == "quotient_value":
                pf.quotient.values[0][0] ^= 1
            elif what == "fri_value":
                pf.fri[0].values[0][0] ^= 1
            elif what == "commitment":
                pf.trace_root[0] ^= 1
            elif what == "fri_layers":
                pf.fri_roots.pop()
            elif what == "zero_commitment":
                pf.trace_root = [0] * 8
            elif what == "last_layer":
                pf.last_layer[0][1] ^= 1
            elif what == "query":
                pf.queries[3] ^= 1
            elif what == "path":
                pf.quotient.paths[5][2][0] ^= 1
        bad = _tampered(proof, mutate)
    assert bad != proof
    assert oracle.verify(bad) != 0


def test_determinism_and_sensitivity(oracle):
    # prover.rs:2300-2330
    p1 = oracle.prove(1, 1, 6)
    assert p1 == oracle.prove(1, 1, 6)
    p2 = oracle.prove(1, 2, 6)
    assert krak.parse(p1).trace_root != krak.parse(p2).trace_root


def test_golden_fixtures(oracle):
    with open(os.path.join(GOLDEN, "proofs.json")) as f:
        cases = json.load(f)
    assert cases
    for c in cases:
        if c["log_n"] > 12:
            continue  # large cases are checked by the GPU tests
        proof = oracle.prove(c["a"], c["b"], c["log_n"], c["n_queries"], c["log_last"])
        assert hashlib.sha256(proof).hexdigest() == c["sha256"], c
        pf = krak.parse(proof)
        assert pf.trace_root == c["trace_root"]
        assert pf.quotient_root == c["quotient_root"]
        assert pf.fri_roots == c["fri_roots"]
        assert pf.queries == c["queries"]
