"""Writes tests/golden/proofs.json: golden proof summaries produced by the oracle.

Every case records the SHA-256 of the KRAK v1 bytes and the commitments,
queries and last layer, plus the first/last 64 LDE values. The oracle itself
is pinned by tests/test_oracle_*.py (hashlib, integer arithmetic and an
independent Python model); these fixtures freeze its output so the GPU tests
can compare against it on a machine without the oracle's runtime budget.

Cases: the reference verifier grid (verifier.rs:369-384), log_n = 10, the
reference-protocol and configured runs of the plain CPU workload (one
Fibonacci column, 2^16 rows, a/b from splitmix64 seed 1), and the plain GPU
workload (2^24 rows, a/b from splitmix64 seed 2, 100 queries, last layer 2^4).

tests/golden/large.json holds commitment summaries (trace and quotient roots,
FRI roots, last layer, query indices) of the largest proofs, 2^26, 2^28 and
2^29 rows (the last one evaluated on the whole circle group, subgroup(31)),
made by the oracle's memory-lean summary path (orc_prove_summary, pinned
against full proofs by tests/test_oracle_large.py): full trees of those sizes
do not fit host memory. a/b from splitmix64 seeds 26, 28, 29; 80 queries,
last layer 2^3.

tests/golden/wide.json holds the batched wide commitment of 34 uniform M31
columns (4 x 2^14, 10 x 2^18, 20 x 2^20 from splitmix64 seed 3, blowup 4,
last layer 2^4): lifted root, FRI roots, last layer, beta and the SHA-256 of
every LDE column.

    python tests/golden/make_golden.py [--skip-large] [--wide-only] [--large-only]
"""

from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import inputs  # noqa: E402
import krak  # noqa: E402
import oracle_lib  # noqa: E402

P = (1 << 31) - 1


def seeded_ab(seed: int):
    a, b = oracle_lib.splitmix64_stream(seed, 2)
    return a % P, b % P


def case(a, b, log_n, nq, last, label):
    t = time.time()
    proof = oracle_lib.prove(a, b, log_n, nq, last)
    assert oracle_lib.verify(proof, nq, last) == 0, label
    pf = krak.parse(proof)
    ev = oracle_lib.lde(oracle_lib.fibonacci_trace(a, b, log_n), log_n)
    print(f"{label}: log_n={log_n} {len(proof)} bytes {time.time() - t:.1f}s", flush=True)
    return {
        "label": label, "a": a, "b": b, "log_n": log_n, "n_queries": nq, "log_last": last,
        "bytes": len(proof), "sha256": hashlib.sha256(proof).hexdigest(),
        "trace_root": pf.trace_root, "quotient_root": pf.quotient_root, "fri_roots": pf.fri_roots,
        "last_layer": pf.last_layer, "queries": pf.queries,
        "lde_head": [int(x) for x in ev[:64]], "lde_tail": [int(x) for x in ev[-64:]],
        "lde_sha256": hashlib.sha256(ev.tobytes()).hexdigest(),
    }


def wide_case(label, seed, log_sizes, log_blowup, log_last):
    t = time.time()
    cols = inputs.m31_columns(seed, log_sizes)
    root, fri, last, beta = oracle_lib.wide_commit(cols, log_sizes, log_blowup, log_last)
    ldes = [hashlib.sha256(oracle_lib.stwo_lde(c, l, log_blowup).tobytes()).hexdigest()
            for c, l in zip(cols, log_sizes)]
    print(f"{label}: {len(cols)} columns, {time.time() - t:.1f}s", flush=True)
    return {"label": label, "seed": seed, "log_sizes": list(log_sizes), "log_blowup": log_blowup,
            "log_last": log_last, "root": [int(x) for x in root], "fri_roots": fri.tolist(),
            "last_layer": last.tolist(), "beta": [int(x) for x in beta], "lde_sha256": ldes}


def write_wide():
    cases = [wide_case("wide_config", inputs.WIDE_SEED, inputs.WIDE_LOG_SIZES, 2, 4),
             wide_case("wide_small", 12, [5, 5, 7, 8, 8, 8], 2, 2)]
    with open(os.path.join(HERE, "wide.json"), "w") as f:
        json.dump(cases, f, separators=(",", ":"))


def summary_case(seed: int, log_n: int, nq: int, last: int):
    t = time.time()
    a, b = seeded_ab(seed)
    s = oracle_lib.prove_summary(a, b, log_n, nq, last)
    print(f"summary log_n={log_n}: {time.time() - t:.1f}s", flush=True)
    return {"label": f"summary_log{log_n}", "a": a, "b": b, "log_n": log_n, "n_queries": nq, "log_last": last, **s}


def write_large():
    cases = [summary_case(n, n, 80, 3) for n in (26, 28, 29)]
    with open(os.path.join(HERE, "large.json"), "w") as f:
        json.dump(cases, f, separators=(",", ":"))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--skip-large", action="store_true")
    ap.add_argument("--wide-only", action="store_true")
    ap.add_argument("--large-only", action="store_true")
    args = ap.parse_args()
    if args.large_only:
        write_large()
        return
    write_wide()
    if args.wide_only:
        return
    cases = []
    for log_n in (4, 5, 6, 8):
        for a, b in ((1, 1), (2, 3), (42, 99), (1000, 2000)):
            cases.append(case(a, b, log_n, 80, 3, f"grid_{a}_{b}_{log_n}"))
    cases.append(case(1, 1, 10, 80, 3, "log10"))
    a, b = seeded_ab(1)
    cases.append(case(a, b, 16, 80, 3, "cpu_plain_reference_protocol"))
    cases.append(case(a, b, 16, 64, 4, "cpu_plain"))
    if not args.skip_large:
        a, b = seeded_ab(2)
        cases.append(case(a, b, 24, 100, 4, "gpu_plain"))
    with open(os.path.join(HERE, "proofs.json"), "w") as f:
        json.dump(cases, f, separators=(",", ":"))


if __name__ == "__main__":
    main()
