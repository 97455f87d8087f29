"""Pins of the oracle pieces used at the largest sizes (CPU only): the batch
inverse's zero convention, the whole-group domain subgroup(k) that evaluation
switches to at 2^31 points, the layer-from-layer twiddle builder and the
memory-lean proof summary (against the full prover, which is pinned by the
other oracle tests)."""

import numpy as np
import pytest

import krak
import pyref

P = (1 << 31) - 1


def reference_batch_inverse(x):
    """The rule cuda/batch_inverse.cu:8-41 implements, read from the kernel: one
    thread per 64-word chunk forms prefix products that skip zero words but
    start from the chunk's first word as it is. So zeros map to zero and every
    other word to its inverse, except that a chunk whose first word is zero has
    a zero product, and every word of it maps to zero."""
    out = []
    for i, v in enumerate(x):
        first = x[i - i % 64]
        out.append(0 if v == 0 or first == 0 else pow(v, P - 2, P))
    return out


def test_batch_inverse_zero_convention(oracle):
    """Zeros sit first, last, inside and alone in chunks, and fill whole
    chunks; the oracle's restatement gives the kernel's rule word for word
    (including the chunk a leading zero clears)."""
    rng = np.random.default_rng(7)
    n = 64 * 9 + 17
    x = rng.integers(1, P, n, dtype=np.uint64).astype(np.uint32)
    zeros = [1, 63, 64, 100, 127, 128 + 5, 130, 131, 64 * 9 + 16]
    zeros += list(range(192, 256))  # a whole chunk
    zeros += [i for i in range(256, 320) if i != 300]  # a chunk with one non-zero word
    zeros += [448]  # a leading zero: its chunk maps to zero
    x[zeros] = 0
    out = np.zeros(n, np.uint32)
    oracle.lib().orc_batch_inverse(x.ctypes.data, out.ctypes.data, n)
    exp = reference_batch_inverse(x.tolist())
    assert out.tolist() == exp
    assert all(w == 0 for w in exp[448:512]) and all(w == 0 for w in exp[64:128])
    assert exp[129] * int(x[129]) % P == 1
    nz = [i for i in range(n) if x[i] and x[i - i % 64]]
    assert all(int(x[i]) * int(out[i]) % P == 1 for i in nz)


@pytest.mark.parametrize("k", [2, 3, 5, 9, 12])
def test_subgroup_twiddles_against_exponents(oracle, k):
    """NTT twiddles on subgroup(k) (the evaluation domain at 2^31 points,
    prover.rs:626-630): circle[i] = y(G^(i 2^(31-k))); line layer l entry i =
    x(G^(2i 4^l 2^(31-k))) (initial (1, 0), so its doublings stay (1, 0))."""
    n = 1 << k
    unit = 1 << (31 - k)
    line, circle, _, _, offs, sizes = oracle.ntt_twiddles(0, k)
    lline, lcircle, _, _ = oracle.ntt_twiddles_lean(0, k)
    assert np.array_equal(line, lline) and np.array_equal(circle, lcircle)
    idx = range(n // 2) if k <= 9 else [0, 1, 2, 3, n // 4, n // 2 - 1]
    for i in idx:
        assert circle[i] == pyref.gpow(i * unit)[1]
    for lyr in range(k - 1):
        assert offs[lyr] + sizes[lyr] <= n - 1 and sizes[lyr] == n >> (lyr + 1)
        for i in (range(sizes[lyr]) if k <= 9 else [0, 1, sizes[lyr] - 1]):
            assert line[offs[lyr] + i] == pyref.gpow(2 * i * (4 ** lyr) * unit)[0], (lyr, i)


@pytest.mark.parametrize("k", [2, 5, 8, 11])
def test_subgroup_circle_fold_twiddles(oracle, k):
    """Circle-fold twiddles on subgroup(k) (prover.rs:754-757, the fold domain at
    2^31 points): the batch inverse of y(G^(bitrev(2i) 2^(31-k))). Only i = 0,
    the point (1, 0), has y = 0 (bitrev(2i) < 2^(k-1) never reaches (-1, 0));
    it leads the first 64-word chunk, so the batch inverse clears that whole
    chunk."""
    tw = oracle.fold_twiddles(0, k, True)
    unit = 1 << (31 - k)
    ys = [pyref.gpow(pyref.bitrev(2 * i, k) * unit)[1] for i in range(1 << (k - 1))]
    assert ys[0] == 0 and all(ys[1:])
    assert tw.tolist() == reference_batch_inverse(ys)
    assert np.all(tw[:64] == 0) and np.all(tw[64:] != 0)


@pytest.mark.parametrize("kind", [0, 1, 2])
@pytest.mark.parametrize("k", [1, 2, 3, 6, 10, 15])
def test_lean_twiddles_equal_full(oracle, kind, k):
    full = oracle.ntt_twiddles(kind, k)[:4]
    lean = oracle.ntt_twiddles_lean(kind, k)
    for a, b in zip(full, lean):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("log_n,nq,last", [(4, 80, 3), (5, 80, 3), (8, 64, 4), (12, 100, 4), (15, 80, 3)])
def test_summary_matches_full_proof(oracle, log_n, nq, last):
    a, b = 1000 + log_n, 77 * log_n
    pf = krak.parse(oracle.prove(a, b, log_n, nq, last))
    s = oracle.prove_summary(a, b, log_n, nq, last)
    assert s["trace_root"] == pf.trace_root
    assert s["quotient_root"] == pf.quotient_root
    assert s["fri_roots"] == pf.fri_roots
    assert s["last_layer"] == pf.last_layer
    assert s["queries"] == pf.queries


def test_summary_parameters(oracle):
    with pytest.raises(AssertionError):
        oracle.prove_summary(1, 2, 30, 80, 3)  # 2^32 points: beyond the circle group
