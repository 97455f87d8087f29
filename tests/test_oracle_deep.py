"""Pins of the oracle's stwo DEEP-quotient pieces (cuda/quotient.cu,
cuda/accumulate_lift.cu; callers quotient_ops.rs, accumulation_ops.rs) by the
mathematics they implement, in Python integers (pyref's QM31 product is a
generic polynomial reduction, independent of the oracle):

- the lifted index ((row >> (r + 1)) << 1) | (row & 1) is the row of the
  circle-group point doubled r times (checked point by point);
- the combined quotient equals sum_j numer_j(q) * (-2u) / V_j(q), with V_j the
  line through the sample point P_j and its conjugate evaluated at the domain
  point q (the CM31 denominator of quotient.cu is V_j / (-2u));
- with true out-of-domain sample values of low-degree columns (some lifted) and
  stwo-style line coefficients, the quotient is low degree: the top half of its
  stwo coefficients vanishes;
- lift_and_accumulate's chain equals one sum at composed lifts.
The reference's tests compare these against stwo's CPU backend at run time and
ship no vectors; stwo is absent here."""

import numpy as np
import pytest

import oracle_lib as O
import pyref as R

P = R.P


def qm(x):
    return tuple(int(v) % P for v in x)


def qinv(x):
    r, b, e = (1, 0, 0, 0), x, P ** 4 - 2
    while e:
        if e & 1:
            r = R.qmul(r, b)
        b = R.qmul(b, b)
        e >>= 1
    return r


def conj(x):  # a + b u -> a - b u
    return (x[0], x[1], (-x[2]) % P, (-x[3]) % P)


def emb(a):
    return (a % P, 0, 0, 0)


def rand_q(rng):
    return tuple(int(v) for v in rng.integers(0, P, 4))


def qpoint(rng):
    """Random QM31 point on the circle: x = (1 - t^2) / (1 + t^2), y = 2t / (1 + t^2)."""
    t = rand_q(rng)
    t2 = R.qmul(t, t)
    d = qinv(R.qadd((1, 0, 0, 0), t2))
    return R.qmul(R.qsub((1, 0, 0, 0), t2), d), R.qmul(R.qadd(t, t), d)


def qdouble(pt):
    x, y = pt
    return R.qsub(R.qscale(R.qmul(x, x), 2), (1, 0, 0, 0)), R.qscale(R.qmul(x, y), 2)


def eval_q(coeffs, pt):
    """stwo circle polynomial (M31 coefficients) at a QM31 point, by the folding of
    pyref.eval_circle_poly (oods.rs:133-215) in QM31 arithmetic."""
    k = len(coeffs).bit_length() - 1
    x, y = pt
    raw = [y]
    for _ in range(1, k):
        raw.append(x)
        x = R.qsub(R.qscale(R.qmul(x, x), 2), (1, 0, 0, 0))
    fs = raw[::-1]

    def fold(vals, f):
        if len(vals) == 1:
            return emb(vals[0])
        h = len(vals) // 2
        return R.qadd(fold(vals[:h], f[1:]), R.qmul(f[0], fold(vals[h:], f[1:])))

    return fold(list(coeffs), fs)


def lift(row, r):
    return ((row >> (r + 1)) << 1) | (row & 1)


def test_lift_index_is_group_doubling():
    L = 7
    for l in range(1, L + 1):
        r = L - l
        for row in range(1 << L):
            q = R.stwo_point(L, R.bitrev(row, L))
            for _ in range(r):
                q = R.pmul(q, q)
            assert q == R.stwo_point(l, R.bitrev(lift(row, r), l)), (l, row)


def test_circle_domain_points():
    for log_n in (1, 2, 5, 8):
        xs, ys = O.circle_domain_brev(log_n)
        for row in range(1 << log_n):
            assert (int(xs[row]), int(ys[row])) == R.stwo_point(log_n, R.bitrev(row, log_n))


def build_case(seed, true_values=True):
    """Columns with trace logs 3, 4, 5 (two each) at blowup 1, lifted to the 2^6 domain;
    two sample points, a batch per (point, size). Returns everything both checks need."""
    rng = np.random.default_rng(seed)
    L = 6
    logs = [3, 3, 4, 4, 5, 5]
    coeffs, evals = [], []
    for l in logs:
        c = rng.integers(0, P, 1 << l).astype(np.uint32)
        coeffs.append([int(v) for v in c])
        etw, _ = O.stwo_twiddles(l + 1)
        d = np.zeros(1 << (l + 1), np.uint32)
        d[: 1 << l] = c  # coefficients, zero-extended: bit-reversed values on the 2^(l+1) domain
        O.lib().orc_stwo_ntt_evaluate(d.ctypes.data, etw.ctypes.data, l + 1)
        evals.append(d)
    points = [qpoint(rng), qpoint(rng)]
    alpha = rand_q(rng)
    batches, apow = [], (1, 0, 0, 0)
    for pt in points:
        for size in (4, 5, 6):
            cols = [i for i, l in enumerate(logs) if l + 1 == size]
            r = L - size
            sp = pt
            for _ in range(r):
                sp = qdouble(sp)
            entries = []
            for i in cols:
                v = eval_q(coeffs[i], sp) if true_values else rand_q(rng)
                a0 = R.qsub(conj(v), v)
                c0 = R.qsub(conj(pt[1]), pt[1])
                b0 = R.qsub(R.qmul(v, c0), R.qmul(a0, pt[1]))
                entries.append((i, R.qmul(apow, a0), R.qmul(apow, b0), R.qmul(apow, c0)))
                apow = R.qmul(apow, alpha)
            batches.append((pt, size, entries))
    return L, evals, batches


def run_oracle(L, evals, batches):
    numers, spx, spy, fla, acc_logs = [], [], [], [], []
    for pt, size, entries in batches:
        idx = [i for i, *_ in entries]
        b = [w for _, _, bb, _ in entries for w in bb]
        c = [w for _, _, _, cc in entries for w in cc]
        numers.append(O.accumulate_numerators(evals, idx, b, c, 1 << size))
        a_acc = (0, 0, 0, 0)
        for _, a, _, _ in entries:
            a_acc = R.qadd(a_acc, a)
        spx += list(pt[0])
        spy += list(pt[1])
        fla += list(a_acc)
        acc_logs.append(size)
    dxs, dys = O.circle_domain_brev(L)
    return O.quotients_combine(spx, spy, fla, numers, acc_logs, dxs, dys, L)


def test_quotient_is_numerator_over_pair_vanishing():
    L, evals, batches = build_case(1, true_values=False)
    got = run_oracle(L, evals, batches)
    m2u = (0, 0, P - 2, 0)  # -2u
    for row in range(1 << L):
        q = R.stwo_point(L, R.bitrev(row, L))
        acc = (0, 0, 0, 0)
        for pt, size, entries in batches:
            # the column value at q: the row of the size-2^size domain holding q doubled L - size times
            d = q
            for _ in range(L - size):
                d = R.pmul(d, d)
            srow = next(j for j in range(1 << size) if R.stwo_point(size, R.bitrev(j, size)) == d)
            num = (0, 0, 0, 0)
            for i, a, b, c in entries:
                num = R.qadd(num, R.qsub(R.qmul(c, emb(int(evals[i][srow]))), R.qadd(R.qscale(a, q[1]), b)))
            px, py = pt
            cx, cy = conj(px), conj(py)
            # line through P and conj(P) at q: (Py - cPy) qx + (cPx - Px) qy + (Px cPy - Py cPx)
            v = R.qadd(R.qadd(R.qscale(R.qsub(py, cy), q[0]), R.qscale(R.qsub(cx, px), q[1])),
                       R.qsub(R.qmul(px, cy), R.qmul(py, cx)))
            acc = R.qadd(acc, R.qmul(R.qmul(num, m2u), qinv(v)))
        assert tuple(int(got[k][row]) for k in range(4)) == acc, row


@pytest.mark.parametrize("seed", [2, 3])
def test_true_samples_give_a_low_degree_quotient(seed):
    L, evals, batches = build_case(seed, true_values=True)
    got = run_oracle(L, evals, batches)
    _, itw = O.stwo_twiddles(L)
    for k in range(4):
        c = got[k].copy()
        O.lib().orc_stwo_ntt_interpolate(c.ctypes.data, itw.ctypes.data, L)
        assert not c[1 << (L - 1):].any(), k
        assert c[: 1 << (L - 1)].any()
    # a wrong sample value breaks it
    pt, size, entries = batches[0]
    i, a, b, c = entries[0]
    batches[0] = (pt, size, [(i, a, R.qadd(b, (1, 0, 0, 0)), c)] + entries[1:])
    bad = run_oracle(L, evals, batches)[0].copy()
    O.lib().orc_stwo_ntt_interpolate(bad.ctypes.data, itw.ctypes.data, L)
    assert bad[1 << (L - 1):].any()


def test_lift_and_accumulate_chain_is_one_sum():
    rng = np.random.default_rng(9)
    logs = [1, 2, 2, 4, 5, 7]
    cols = [rng.integers(0, P, (1 << l, 4)).astype(np.uint32) for l in logs]
    curr, curr_log = np.zeros((2, 4), np.uint32), 1
    for col, l in zip(cols, logs):
        curr = O.accumulate_lift(col, curr, l - curr_log).reshape(-1, 4)
        curr_log = l
    L = logs[-1]
    exp = np.zeros((1 << L, 4), np.int64)
    for col, l in zip(cols, logs):
        exp += col[[lift(i, L - l) for i in range(1 << L)]]
    assert np.array_equal(curr, (exp % P).astype(np.uint32))
