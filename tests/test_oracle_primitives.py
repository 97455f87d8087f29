"""Pins of the oracle's primitives against independent computations.

Blake2s against hashlib (RFC 7693); M31/CM31/QM31 against Python integers,
the reference's own known answers and group-order facts; circle points
against exponent arithmetic; the native NTT against its closed-form basis
product; FRI folds against the low-degree behaviour of folding.
"""

import ctypes as C
import hashlib
import random
import struct

import numpy as np
import pytest

import pyref

P = pyref.P


def test_blake2s_matches_hashlib(oracle):
    L = oracle.lib()
    rng = random.Random(1)
    for ln in [0, 1, 3, 4, 16, 37, 63, 64, 65, 127, 128, 129, 160, 288, 1000]:
        msg = bytes(rng.randrange(256) for _ in range(ln))
        out = (C.c_uint8 * 32)()
        L.orc_blake2s(msg, ln, out)
        assert bytes(out) == hashlib.blake2s(msg).digest(), ln
        if ln <= 64:  # the reference's one-block helper equals standard Blake2s up to 64 bytes
            L.orc_blake2s_single(msg, ln, out)
            assert bytes(out) == hashlib.blake2s(msg).digest(), ln
    assert hashlib.blake2s(b"abc").hexdigest().startswith("508c5e8c")


def test_m31_against_integers(oracle):
    L = oracle.lib()
    rng = random.Random(2)
    vals = [0, 1, 2, P - 1, P - 2, 1 << 30, (1 << 30) + 1] + [rng.randrange(P) for _ in range(300)]
    for a in vals[:40]:
        for b in vals[-40:]:
            assert L.orc_m31_add(a, b) == (a + b) % P
            assert L.orc_m31_sub(a, b) == (a - b) % P
            assert L.orc_m31_mul(a, b) == (a * b) % P
    for a in vals[1:60]:
        if a:
            assert L.orc_m31_inv(a) * a % P == 1
    # reference known answers (m31.rs:146-178)
    assert L.orc_m31_add(P - 1, 2) == 1
    assert L.orc_m31_sub(0, 1) == P - 1
    assert L.orc_m31_mul(1 << 16, 1 << 16) == (1 << 32) % P


def q(L, fn, a, b=None):
    x = oracle_arr(a)
    out = (C.c_uint32 * 4)()
    if b is None:
        fn(x, out)
    else:
        fn(x, oracle_arr(b), out)
    return tuple(out)


def oracle_arr(v):
    return (C.c_uint32 * len(v))(*v)


def test_cm31_qm31_against_polynomial_arithmetic(oracle):
    L = oracle.lib()
    # cm31.rs:120-129: (3+4i)(5+6i) = -9 + 38i
    out = (C.c_uint32 * 2)()
    L.orc_cm31_mul(oracle_arr([3, 4]), oracle_arr([5, 6]), out)
    assert tuple(out) == (P - 9, 38)
    # u * u = 2 + i
    assert q(L, L.orc_qm31_mul, [0, 0, 1, 0], [0, 0, 1, 0]) == (2, 1, 0, 0)
    rng = random.Random(3)
    for _ in range(200):
        a = [rng.randrange(P) for _ in range(4)]
        b = [rng.randrange(P) for _ in range(4)]
        assert q(L, L.orc_qm31_mul, a, b) == pyref.qmul(a, b)
    # inverse round trips (qm31.rs:190-196) and the group order p^4 - 1
    for v in ([111, 222, 333, 444], [1, 0, 0, 0], [0, 1, 0, 0], [5, 0, 7, 0]):
        iv = q(L, L.orc_qm31_inv, v)
        assert pyref.qmul(v, iv) == (1, 0, 0, 0)
    a = (rng.randrange(P), rng.randrange(P), rng.randrange(P), rng.randrange(P))
    acc, base, e = (1, 0, 0, 0), a, P**4 - 1
    while e:
        if e & 1:
            acc = q(L, L.orc_qm31_mul, list(acc), list(base))
        base = q(L, L.orc_qm31_mul, list(base), list(base))
        e >>= 1
    assert acc == (1, 0, 0, 0)


def test_circle_group_facts(oracle):
    L = oracle.lib()
    g = oracle_arr([2, 1268011823])
    assert (4 + 1268011823**2) % P == 1  # generator on the circle (circle.rs:586-592)
    out = (C.c_uint32 * 2)()
    L.orc_circle_repeated_double(g, 31, out)
    assert tuple(out) == (1, 0)
    L.orc_circle_repeated_double(g, 30, out)
    assert tuple(out) == (P - 1, 0)  # order exactly 2^31
    for kind, expf in ((1, pyref.half_coset_exp), (2, pyref.half_odds_exp)):
        for k in (1, 3, 7, 12, 20):
            for i in (0, 1, 5, (1 << k) - 1):
                L.orc_coset_at(kind, k, i, out)
                assert tuple(out) == pyref.gpow(expf(k, i)), (kind, k, i)
    # subgroup(k).at(i) = G^(i 2^(31-k))
    for k in (2, 5, 10):
        for i in (0, 1, 3):
            L.orc_coset_at(0, k, i, out)
            assert tuple(out) == pyref.gpow(i << (31 - k))


@pytest.mark.parametrize("k", [2, 3, 5, 8, 11])
def test_ntt_twiddle_layout(oracle, k):
    line, circle, iline, icircle, offs, sizes = oracle.ntt_twiddles(1, k)
    T = pyref.ntt_tables(k)
    assert list(circle) == T[0]
    for lyr in range(k - 1):
        assert sizes[lyr] == (1 << k) >> (lyr + 1)
        seg = line[offs[lyr]: offs[lyr] + sizes[lyr]]
        assert list(seg) == T[lyr + 1], lyr
    assert all(int(a) * int(b) % P == 1 for a, b in zip(line[: (1 << k) - 2], iline))
    assert all(int(a) * int(b) % P == 1 for a, b in zip(circle, icircle))


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5, 6])
def test_ntt_matches_closed_form_basis(oracle, k):
    rng = random.Random(10 + k)
    n = 1 << k
    line, circle, iline, icircle, offs, sizes = oracle.ntt_twiddles(1, k)
    coeffs = np.array([rng.randrange(P) for _ in range(n)], np.uint32)
    data = coeffs.copy()
    oracle.lib().orc_ntt_evaluate(data.ctypes.data, line.ctypes.data, circle.ctypes.data, k)
    assert list(data) == pyref.ntt_eval_closed_form([int(c) for c in coeffs], k)
    oracle.lib().orc_ntt_interpolate(data.ctypes.data, iline.ctypes.data, icircle.ctypes.data, k)
    assert np.array_equal(data, coeffs)


@pytest.mark.parametrize("k", [4, 10, 16])
def test_ntt_roundtrip_const_linearity(oracle, k):
    # ntt.rs:349-366 (roundtrip), ntt.rs:694-700 ([1,0,...] -> all ones)
    L = oracle.lib()
    n = 1 << k
    line, circle, iline, icircle, _, _ = oracle.ntt_twiddles(1, k)
    x = np.array([(i * 7 + 13) % P for i in range(n)], np.uint32)
    d = x.copy()
    L.orc_ntt_evaluate(d.ctypes.data, line.ctypes.data, circle.ctypes.data, k)
    L.orc_ntt_interpolate(d.ctypes.data, iline.ctypes.data, icircle.ctypes.data, k)
    assert np.array_equal(d, x)
    one = np.zeros(n, np.uint32)
    one[0] = 1
    L.orc_ntt_evaluate(one.ctypes.data, line.ctypes.data, circle.ctypes.data, k)
    assert np.all(one == 1)
    rng = np.random.default_rng(k)
    a = rng.integers(0, P, n, dtype=np.uint64).astype(np.uint32)
    b = rng.integers(0, P, n, dtype=np.uint64).astype(np.uint32)
    s = ((a.astype(np.uint64) + b) % P).astype(np.uint32)
    for v in (a, b, s):
        L.orc_ntt_evaluate(v.ctypes.data, line.ctypes.data, circle.ctypes.data, k)
    assert np.array_equal(((a.astype(np.uint64) + b) % P).astype(np.uint32), s)


def test_fold_twiddles_against_exponents(oracle):
    for k in (2, 3, 6, 9):
        assert list(oracle.fold_twiddles(1, k, True)) == pyref.fold_twiddles_circle(k)
        assert list(oracle.fold_twiddles(2, k, False)) == pyref.fold_twiddles_line(k)


def test_line_fold_halves_degree(oracle):
    """A line polynomial of degree < 2^d on half_odds(k) in bit-reversed order folds
    to degree < 2^(d-1); after d folds it is constant. Pins the fold formula,
    its twiddles and the pairing (fri.rs:411-427, twiddles.cu:23-52)."""
    rng = random.Random(7)
    k, d = 7, 3
    n = 1 << k
    # basis 1, x, pi(x), x pi(x), ... with pi(x) = 2x^2 - 1 (fri.rs:323-331)
    coeffs = [[rng.randrange(P) for _ in range(4)] for _ in range(1 << d)]

    def line_eval(x):
        acc = [0, 0, 0, 0]
        for idx, c in enumerate(coeffs):
            t, y, m = 1, x, idx
            while m:
                if m & 1:
                    t = t * y % P
                y = (2 * y * y - 1) % P
                m >>= 1
            acc = [(acc[i] + c[i] * t) % P for i in range(4)]
        return acc

    vals = [line_eval(pyref.gpow(pyref.half_odds_exp(k, pyref.bitrev(j, k)))[0]) for j in range(n)]
    cols = np.array(vals, np.uint32).T.copy()
    for step in range(d):
        tw = oracle.fold_twiddles(2, k - step, False)
        alpha = [rng.randrange(P) for _ in range(4)]
        cols = oracle.fold_line(cols, tw, alpha)
    assert np.all(cols == cols[:, :1]), "folded layer is not constant"


def test_merkle_root_against_hashlib(oracle):
    rng = np.random.default_rng(5)
    for n_cols, log_n in ((1, 0), (1, 5), (4, 6), (2, 4), (16, 3), (20, 3)):
        cols = rng.integers(0, P, (n_cols, 1 << log_n), dtype=np.uint64).astype(np.uint32)
        eff = min(n_cols, 16)
        leaves = [pyref.leaf_hash([int(cols[c][i]) for c in range(eff)]) for i in range(1 << log_n)]
        root = pyref.merkle_layers(leaves)[-1][0]
        assert list(oracle.merkle_root(cols)) == root
    # auth path verification
    cols = rng.integers(0, P, (1, 64), dtype=np.uint64).astype(np.uint32)
    leaves = [pyref.leaf_hash([int(v)]) for v in cols[0]]
    layers = pyref.merkle_layers(leaves)
    root = np.array(layers[-1][0], np.uint32)
    for idx in (0, 17, 63):
        path = np.array([w for h in pyref.auth_path(layers, idx) for w in h], np.uint32)
        leaf = np.array(leaves[idx], np.uint32)
        L = oracle.lib()
        assert L.orc_merkle_verify_path(oracle.p32(root), oracle.p32(leaf), idx, oracle.p32(path), 6) == 1
        assert L.orc_merkle_verify_path(oracle.p32(root), oracle.p32(leaf), idx ^ 1, oracle.p32(path), 6) == 0


def test_channel_against_hashlib(oracle):
    L = oracle.lib()
    # channel.rs:346-381: mix_felts of 9 QM31 (1..36) from the zero state
    ch = oracle.ChannelStruct()
    L.orc_channel_init(C.byref(ch))
    words = oracle_arr(list(range(1, 37)))
    L.orc_channel_mix_felts(C.byref(ch), words, 9)
    assert bytes(ch.state) == hashlib.blake2s(bytes(32) + struct.pack("<36I", *range(1, 37))).digest()
    # transcript draws against the hashlib model
    ref = pyref.Channel()
    L.orc_channel_init(C.byref(ch))
    for step in range(6):
        d = [random.Random(step).randrange(1 << 32) for _ in range(8)]
        L.orc_channel_mix_digest(C.byref(ch), oracle_arr(d))
        ref.mix_digest(d)
        out = (C.c_uint32 * 4)()
        L.orc_channel_draw_felt(C.byref(ch), out)
        assert tuple(out) == ref.draw_felt()
        for bound in (7, 1 << 18, 1 << 26):
            assert L.orc_channel_draw_number(C.byref(ch), bound) == ref.draw_number(bound)
    assert bytes(ch.state) == ref.state


def test_quotient_and_trace(oracle):
    t = oracle.fibonacci_trace(3, 5, 6)
    assert t[0] == 3 and t[1] == 5
    assert all((int(t[i]) + int(t[i + 1])) % P == t[i + 2] for i in range(62))
    rng = random.Random(9)
    ev = [rng.randrange(P) for _ in range(32)]
    alpha = [rng.randrange(P) for _ in range(4)]
    qs = oracle.quotient(ev, alpha)
    for i in range(32):
        c = (ev[(i + 2) % 32] - ev[(i + 1) % 32] - ev[i]) % P
        assert tuple(int(x) for x in qs[:, i]) == pyref.qscale(alpha, c)
