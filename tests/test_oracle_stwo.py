"""Pins of the oracle's stwo-basis NTT, lifted leaves and wide commitment.

The stwo NTT is checked against direct evaluation of the circle polynomial by
recursive folding (src/oods.rs:133-215) at the canonic domain points in
bit-reversed order — the check the reference itself makes at ntt.rs:480-513.
"""

import ctypes as C
import random

import numpy as np
import pytest

import pyref

P = pyref.P


def oracle_stwo(oracle, log_n):
    n = 1 << log_n
    f = np.zeros(max(1, n // 2), np.uint32)
    i = np.zeros_like(f)
    oracle.lib().orc_stwo_twiddles(log_n, f.ctypes.data, i.ctypes.data)
    return f, i


@pytest.mark.parametrize("k", [2, 3, 5, 8, 11])
def test_stwo_twiddles_against_exponents(oracle, k):
    f, i = oracle_stwo(oracle, k)
    assert list(f) == pyref.stwo_twiddles(k)
    assert all(int(a) * int(b) % P == 1 for a, b in zip(f, i))


@pytest.mark.parametrize("k", [3, 4, 6, 8])
def test_stwo_evaluate_is_circle_poly_evaluation(oracle, k):
    rng = random.Random(k)
    n = 1 << k
    f, i = oracle_stwo(oracle, k)
    coeffs = [rng.randrange(P) for _ in range(n)]
    d = np.array(coeffs, np.uint32)
    oracle.lib().orc_stwo_ntt_evaluate(d.ctypes.data, f.ctypes.data, k)
    assert list(d) == pyref.stwo_evaluate(coeffs, k)
    oracle.lib().orc_stwo_ntt_interpolate(d.ctypes.data, i.ctypes.data, k)
    assert list(d) == coeffs


def test_stwo_size4_reads_pad_word(oracle):
    # At n = 4 the first line layer holds one word, so the circle layer pairs it with the
    # pad word 1 (circle_ntt_stwo.cu:202-231 reads first_line_layer[0..2)): circle
    # twiddles are [1, -1, -x, x] with x = 32768, not the domain's y coordinates.
    f, i = oracle_stwo(oracle, 2)
    assert list(f) == [32768, 1]
    d = np.array([0, 1, 0, 0], np.uint32)
    oracle.lib().orc_stwo_ntt_evaluate(d.ctypes.data, f.ctypes.data, 2)
    assert list(d) == [1, P - 1, P - 1, 1]
    d = np.array([0, 0, 0, 1], np.uint32)
    oracle.lib().orc_stwo_ntt_evaluate(d.ctypes.data, f.ctypes.data, 2)
    # line layer: v1 = 32768, v3 = -32768; circle layer t = 1 on (v0, v1), -1 on (v2, v3)
    assert list(d) == [32768, P - 32768, 32768, P - 32768]


def test_stwo_reference_known_coefficients(oracle):
    # ntt.rs:516-530: coefficients [42, 17, 5, 99, 0...] at log 6 and 8
    for k in (6, 8):
        n = 1 << k
        f, _ = oracle_stwo(oracle, k)
        coeffs = [42, 17, 5, 99] + [0] * (n - 4)
        d = np.array(coeffs, np.uint32)
        oracle.lib().orc_stwo_ntt_evaluate(d.ctypes.data, f.ctypes.data, k)
        for j in range(n):
            pt = pyref.stwo_point(k, pyref.bitrev(j, k))
            assert d[j] == pyref.eval_circle_poly(coeffs, pt)


def test_lifted_leaves_against_hashlib(oracle):
    rng = np.random.default_rng(3)
    # the reference's mixed shape: 38 columns over 4 sizes (tests.rs:208-254), scaled down
    logs = [3] * 5 + [4] * 9 + [5] * 17 + [6] * 7
    S = max(logs)
    cols = [rng.integers(0, P, 1 << l, dtype=np.uint64).astype(np.uint32) for l in logs]
    ptrs = (C.c_void_p * len(cols))(*[c.ctypes.data for c in cols])
    lg = np.array(logs, np.uint32)
    out = np.zeros(8 << S, np.uint32)
    oracle.lib().orc_lifted_leaves(ptrs, lg.ctypes.data, len(cols), S, out.ctypes.data)
    for leaf in (0, 1, 2, 17, (1 << S) - 1):
        assert list(out[8 * leaf: 8 * leaf + 8]) == pyref.lifted_leaf(cols, logs, S, leaf)
    # lifted row formula (merkle_leaves_lifted.cu:26-31)
    for leaf in range(64):
        for L in (3, 4, 6):
            assert oracle.lib().orc_lifted_row(leaf, 6, L) == pyref.lifted_row(leaf, 6, L)


def test_pack_and_bit_reverse(oracle):
    rng = np.random.default_rng(4)
    n = 64
    ins = [rng.integers(0, P, n, dtype=np.uint64).astype(np.uint32) for _ in range(4)]
    outs = [np.zeros(n // 16, np.uint32) for _ in range(64)]
    ip = (C.c_void_p * 4)(*[a.ctypes.data for a in ins])
    op = (C.c_void_p * 64)(*[a.ctypes.data for a in outs])
    oracle.lib().orc_pack_leaves(ip, op, n)
    for c in range(4):
        for j in range(n):
            assert outs[c * 16 + j % 16][j // 16] == ins[c][j]
    q = rng.integers(0, P, 4 * 32, dtype=np.uint64).astype(np.uint32)
    r = np.zeros_like(q)
    oracle.lib().orc_bit_reverse_qm31(q.ctypes.data, r.ctypes.data, 32, 5)
    for i in range(32):
        j = pyref.bitrev(i, 5)
        assert list(r[4 * j: 4 * j + 4]) == list(q[4 * i: 4 * i + 4])


def wide_oracle(oracle, cols, logs, blowup, log_last):
    ptrs = (C.c_void_p * len(cols))(*[c.ctypes.data for c in cols])
    lg = np.array(logs, np.uint32)
    root = np.zeros(8, np.uint32)
    S = max(logs) + blowup
    fri = np.zeros(8 * 64, np.uint32)
    last = np.zeros(4 << log_last, np.uint32)
    beta = np.zeros(4, np.uint32)
    nf = oracle.lib().orc_wide_commit(ptrs, lg.ctypes.data, len(cols), blowup, log_last, oracle.p32(root),
                                      fri.ctypes.data, last.ctypes.data, beta.ctypes.data)
    assert nf == S - log_last
    return root, fri[: 8 * nf].reshape(-1, 8), last, beta


def test_wide_commit_against_python_model(oracle):
    rng = np.random.default_rng(5)
    logs = [3, 3, 4, 5, 5]
    cols = [rng.integers(0, P, 1 << l, dtype=np.uint64).astype(np.uint32) for l in logs]
    root, fri, last, beta = wide_oracle(oracle, cols, logs, 2, 2)
    r2, roots2, last2, beta2 = pyref.wide_commit([[int(v) for v in c] for c in cols], logs, 2, 2)
    assert list(root) == r2
    assert [list(r) for r in fri] == roots2
    assert list(last) == last2
    assert tuple(beta) == beta2


def test_splitmix_inputs_match_oracle_generator(oracle):
    import inputs

    ref = oracle.splitmix64_stream(7, 40)
    assert [int(x) for x in inputs.splitmix64(7, 40)] == ref
    assert [int(x) for x in inputs.splitmix64(7, 10, 30)] == ref[30:]
    cols = inputs.m31_columns(9, [2, 3])
    s = oracle.splitmix64_stream(9, 12)
    assert [int(x) for x in cols[0]] == [v % P for v in s[:4]]
    assert [int(x) for x in cols[1]] == [v % P for v in s[4:]]


def test_wide_commit_rejects_unsorted_sizes(oracle):
    cols = [np.zeros(16, np.uint32), np.zeros(8, np.uint32)]
    ptrs = (C.c_void_p * 2)(*[c.ctypes.data for c in cols])
    lg = np.array([4, 3], np.uint32)
    root = np.zeros(8, np.uint32)
    assert oracle.lib().orc_wide_commit(ptrs, lg.ctypes.data, 2, 2, 2, oracle.p32(root), None, None, None) == 0
