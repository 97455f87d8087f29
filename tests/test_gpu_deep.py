"""stwo DEEP-quotient kernels on the MI355X against the oracle, bit-exact
(cuda/quotient.cu, cuda/accumulate_lift.cu; ffi.rs:840-897, 957-958): edge
sizes (1 and 2 rows, rows not a multiple of a workgroup's), empty batches,
repeated columns, zero denominators, misaligned QM31 arrays, and the wide
commitment's sizes (2^16, 2^20, 2^22 lifted to 2^22)."""

import ctypes as C

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu
P = (1 << 31) - 1


def rnd(rng, *shape):
    return rng.integers(0, P, shape, dtype=np.uint64).astype(np.uint32)


def dev_ptrs(vx, bufs):
    addrs = np.array([b.addr for b in bufs], dtype=np.uint64)
    d = vx.DeviceBuffer(2 * max(1, len(bufs)))
    if len(bufs):
        vx.lib.vx_memcpy_h2d(d.ptr, addrs.ctypes.data, 8 * len(bufs))
    return d


def gpu_numerators(vx, cols, idx, b, c, n_rows):
    dcols = [vx.DeviceBuffer.from_host(x) for x in cols]
    dp = dev_ptrs(vx, dcols)
    di = vx.DeviceBuffer.from_host(np.asarray(idx, np.uint32) if len(idx) else np.zeros(1, np.uint32))
    db = vx.DeviceBuffer.from_host(np.asarray(b, np.uint32) if len(b) else np.zeros(4, np.uint32))
    dc = vx.DeviceBuffer.from_host(np.asarray(c, np.uint32) if len(c) else np.zeros(4, np.uint32))
    outs = [vx.DeviceBuffer(n_rows) for _ in range(4)]
    vx.lib.cuda_accumulate_numerators(dp.ptr, di.ptr, db.ptr, dc.ptr, len(idx), n_rows, *[o.ptr for o in outs])
    vx.synchronize()
    vx.check("cuda_accumulate_numerators")
    return np.stack([o.to_host() for o in outs])


@pytest.mark.parametrize("n_rows,n_batch", [(1, 1), (2, 0), (255, 2), (257, 3), (512, 3), (1000, 4), (1536, 6),
                                            (1 << 12, 7), (1 << 20, 5)])
def test_accumulate_numerators(vx, n_rows, n_batch):
    rng = np.random.default_rng(n_rows + n_batch)
    cols = [rnd(rng, n_rows) for _ in range(6)]
    idx = rng.integers(0, 6, n_batch).astype(np.uint32)  # repeats allowed
    b, c = rnd(rng, 4 * n_batch), rnd(rng, 4 * n_batch)
    exp = O.accumulate_numerators(cols, idx, b, c, n_rows)
    got = gpu_numerators(vx, cols, idx, b, c, n_rows)
    assert np.array_equal(got, exp)


def gpu_combine(vx, spx, spy, fla, numers, acc_logs, dxs, dys, L, n_rows):
    keep = []
    ptr_bufs = []
    for k in range(4):
        bufs = [vx.DeviceBuffer.from_host(n[k]) for n in numers]
        keep += bufs
        ptr_bufs.append(dev_ptrs(vx, bufs))
    d = [vx.DeviceBuffer.from_host(np.asarray(x, np.uint32)) for x in (spx, spy, fla, acc_logs, dxs, dys)]
    outs = [vx.DeviceBuffer(n_rows) for _ in range(4)]
    vx.lib.cuda_compute_quotients_combine(d[0].ptr, d[1].ptr, d[2].ptr, *[p.ptr for p in ptr_bufs], d[3].ptr,
                                          len(numers), d[4].ptr, d[5].ptr, L, n_rows, *[o.ptr for o in outs])
    vx.synchronize()
    vx.check("cuda_compute_quotients_combine")
    return np.stack([o.to_host() for o in outs])


@pytest.mark.parametrize("L,acc_logs,n_rows", [(1, [1], 2), (3, [1, 3], 8), (6, [6, 4, 2, 6], 64),
                                               (10, [10, 7], 1000), (12, [12, 9, 12], 1 << 12),
                                               (20, [20, 16, 18], 1 << 20)])
def test_quotients_combine(vx, L, acc_logs, n_rows):
    rng = np.random.default_rng(L * 7 + n_rows)
    n = len(acc_logs)
    spx, spy, fla = rnd(rng, 4 * n), rnd(rng, 4 * n), rnd(rng, 4 * n)
    numers = [rnd(rng, 4, 1 << l) for l in acc_logs]
    dxs, dys = O.circle_domain_brev(L)
    dxs, dys = dxs[:n_rows], dys[:n_rows]
    exp = O.quotients_combine(spx, spy, fla, numers, acc_logs, dxs, dys, L)
    got = gpu_combine(vx, spx, spy, fla, numers, acc_logs, dxs, dys, L, n_rows)
    assert np.array_equal(got, exp)


def test_quotients_combine_zero_denominators(vx):
    """A sample point with M31 coordinates makes every denominator 0 (its terms
    vanish, as cm31_inv(0) = 0 gives); a domain point placed on a sample line makes
    one row's denominator 0 among nonzero neighbours."""
    rng = np.random.default_rng(5)
    L = 9
    dxs, dys = O.circle_domain_brev(L)
    spx, spy, fla = rnd(rng, 8), rnd(rng, 8), rnd(rng, 8)
    spx[2:4] = 0  # accumulation 0: Im(px) = Im(py) = 0
    spy[2:4] = 0
    # accumulation 1: choose Re(px), Re(py) so that den vanishes at row 37:
    # (Re(px) - x) Im(py) = (Re(py) - y) Im(px) with Re(px) = x, Re(py) = y
    spx[4:6] = [dxs[37], 0]
    spy[4:6] = [dys[37], 0]
    numers = [rnd(rng, 4, 1 << L), rnd(rng, 4, 1 << (L - 2))]
    exp = O.quotients_combine(spx, spy, fla, numers, [L, L - 2], dxs, dys, L)
    got = gpu_combine(vx, spx, spy, fla, numers, [L, L - 2], dxs, dys, L, 1 << L)
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("misaligned", [False, True])
@pytest.mark.parametrize("col_log,curr_log", [(1, 1), (3, 1), (10, 4), (18, 18), (20, 12)])
def test_accumulate_lift(vx, col_log, curr_log, misaligned):
    rng = np.random.default_rng(col_log * 31 + curr_log)
    col, curr = rnd(rng, 1 << col_log, 4), rnd(rng, 1 << curr_log, 4)
    exp = O.accumulate_lift(col, curr, col_log - curr_log)
    off = 1 if misaligned else 0  # words: 4-byte aligned, not 16-byte aligned
    dcol = vx.DeviceBuffer.from_host(np.concatenate([np.zeros(off, np.uint32), col.ravel()]))
    dcur = vx.DeviceBuffer.from_host(np.concatenate([np.zeros(off, np.uint32), curr.ravel()]))
    vx.lib.cuda_accumulate_lift(C.c_void_p(dcol.addr + 4 * off), C.c_void_p(dcur.addr + 4 * off), 1 << col_log,
                                col_log - curr_log)
    vx.synchronize()
    vx.check("cuda_accumulate_lift")
    assert np.array_equal(dcol.to_host()[off:], exp.ravel())


def test_wide_shape_end_to_end(vx):
    """The wide commitment's LDE sizes (4 x 2^16, 10 x 2^20, 20 x 2^22), two sample
    points, a batch per (point, size): numerators and the combined 2^22-row quotient
    against the oracle."""
    rng = np.random.default_rng(11)
    sizes = [16] * 4 + [20] * 10 + [22] * 20
    L = 22
    batches = []
    for pt in range(2):
        for s in (16, 20, 22):
            cols = [rnd(rng, 1 << s) for i, l in enumerate(sizes) if l == s]
            n = len(cols)
            batches.append((s, cols, rnd(rng, 4 * n), rnd(rng, 4 * n), pt))
    spx_pts, spy_pts = rnd(rng, 2, 4), rnd(rng, 2, 4)
    numers, exp_numers = [], []
    for s, cols, b, c, _ in batches:
        idx = np.arange(len(cols), dtype=np.uint32)
        exp_numers.append(O.accumulate_numerators(cols, idx, b, c, 1 << s))
        numers.append(gpu_numerators(vx, cols, idx, b, c, 1 << s))
        assert np.array_equal(numers[-1], exp_numers[-1]), s
    spx = np.concatenate([spx_pts[pt] for *_, pt in batches])
    spy = np.concatenate([spy_pts[pt] for *_, pt in batches])
    fla = rnd(rng, 4 * len(batches))
    logs = [s for s, *_ in batches]
    dxs, dys = O.circle_domain_brev(L)
    exp = O.quotients_combine(spx, spy, fla, exp_numers, logs, dxs, dys, L)
    got = gpu_combine(vx, spx, spy, fla, numers, logs, dxs, dys, L, 1 << L)
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("log_n", [1, 2, 5, 16, 22])
def test_circle_domain_points(vx, log_n):
    xs, ys = vx.circle_domain_points(log_n)
    vx.synchronize()
    exs, eys = O.circle_domain_brev(log_n)
    assert np.array_equal(xs.to_host(), exs) and np.array_equal(ys.to_host(), eys)


def oracle_pipeline(cols, logs, points, lc, L, dxs, dys):
    """The reference pipeline on the oracle: one accumulation per (point, size)."""
    numers, spx, spy, fla, acc_logs = [], [], [], [], []
    for p in range(points.shape[0]):
        for s in sorted(set(logs)):
            idx = np.array([c for c in range(len(cols)) if logs[c] == s], np.uint32)
            sub = [cols[c] for c in idx]
            b = lc[p, idx, 4:8].ravel()
            c = lc[p, idx, 8:12].ravel()
            numers.append(O.accumulate_numerators(sub, np.arange(len(idx), dtype=np.uint32), b, c, 1 << s))
            a = lc[p, idx, 0:4].astype(np.uint64).sum(axis=0) % P
            spx += list(points[p, :4])
            spy += list(points[p, 4:])
            fla += list(a.astype(np.uint32))
            acc_logs.append(s)
    return O.quotients_combine(spx, spy, fla, numers, acc_logs, dxs, dys, L)


@pytest.mark.parametrize("L,logs,n_points", [(1, [1], 1), (4, [2, 4, 3, 1], 2), (8, [8] * 5 + [5, 6], 3),
                                             (10, [3, 10, 7, 10, 9], 4), (14, [14, 12, 12, 9], 2),
                                             (20, [16] * 2 + [18] * 3 + [20] * 4, 2),
                                             # column counts 2, 3, 6 and 8: no whole group of three, one
                                             # group and no remainder, two groups, two groups and two left
                                             (5, [5, 4], 1), (6, [6, 6, 2], 4), (12, [12, 11, 10, 12, 9, 12], 3),
                                             (9, [9] * 8, 1)])
def test_fused_quotient_matches_reference_pipeline(vx, L, logs, n_points):
    rng = np.random.default_rng(L * 13 + n_points)
    cols = [rnd(rng, 1 << l) for l in logs]
    points = rnd(rng, n_points, 8)
    lc = rnd(rng, n_points, len(cols), 12)
    if n_points > 1:
        lc[1, 0] = 0  # column 0 not sampled at point 1
    dxs, dys = vx.circle_domain_points(L)
    exs, eys = O.circle_domain_brev(L)
    exp = oracle_pipeline(cols, logs, points, lc, L, exs, eys)
    dcols = [vx.DeviceBuffer.from_host(c) for c in cols]
    outs = [vx.DeviceBuffer(1 << L) for _ in range(4)]
    dq = vx.DeepQuotient(len(cols), n_points)
    dq.quotient([c.addr for c in dcols], logs, points, lc, L, dxs.addr, dys.addr, [o.addr for o in outs])
    vx.synchronize()
    vx.check("vx_deep_quotient")
    assert np.array_equal(np.stack([o.to_host() for o in outs]), exp)
    # a second call on the same object, other coefficients: the staging is reused safely
    lc2 = rnd(rng, n_points, len(cols), 12)
    dq.quotient([c.addr for c in dcols], logs, points, lc2, L, dxs.addr, dys.addr, [o.addr for o in outs])
    vx.synchronize()
    assert np.array_equal(np.stack([o.to_host() for o in outs]), oracle_pipeline(cols, logs, points, lc2, L, exs, eys))
    dq.close()


def test_fused_quotient_rejects_bad_shapes(vx):
    dq = vx.DeepQuotient(2, 2)
    buf = vx.DeviceBuffer(16)
    pts, lc = np.zeros((3, 8), np.uint32), np.zeros((3, 2, 12), np.uint32)
    with pytest.raises(vx.HipError):
        dq.quotient([buf.addr, buf.addr], [2, 2], pts, lc, 4, buf.addr, buf.addr, [buf.addr] * 4)  # 3 points > 2
    with pytest.raises(vx.HipError):
        dq.quotient([buf.addr], [5], pts[:1], lc[:1, :1], 4, buf.addr, buf.addr, [buf.addr] * 4)  # log 5 > 4
    dq.close()


def test_fused_quotient_calls_in_flight_on_two_streams(vx):
    """Three calls on one object on two streams with no host wait in between:
    the third call reuses the metadata slot of the first, which ran on the other
    stream, so its stream must wait for that call's kernel."""
    rng = np.random.default_rng(21)
    L, logs, n_points = 12, [12, 10, 12, 11], 2
    cols = [rnd(rng, 1 << l) for l in logs]
    dcols = [vx.DeviceBuffer.from_host(c) for c in cols]
    points = rnd(rng, n_points, 8)
    dxs, dys = vx.circle_domain_points(L)
    exs, eys = O.circle_domain_brev(L)
    vx.synchronize()
    streams = []
    for _ in range(2):
        sp = C.c_void_p()
        assert vx.lib.vx_stream_create(C.byref(sp)) == 0
        streams.append(sp.value)
    dq = vx.DeepQuotient(len(cols), n_points)
    lcs, outs = [], []
    for call in range(3):
        lc = rnd(rng, n_points, len(cols), 12)
        out = [vx.DeviceBuffer(1 << L) for _ in range(4)]
        # calls 0, 1, 2 on streams 0, 1, 1: call 2 reuses call 0's slot from the other stream
        dq.quotient([c.addr for c in dcols], logs, points, lc, L, dxs.addr, dys.addr, [o.addr for o in out],
                    stream=streams[min(call, 1)])
        lcs.append(lc)
        outs.append(out)
    for sp in streams:
        vx.synchronize_stream(sp)
    vx.check("vx_deep_quotient on two streams")
    for lc, out in zip(lcs, outs):
        exp = oracle_pipeline(cols, logs, points, lc, L, exs, eys)
        assert np.array_equal(np.stack([o.to_host() for o in out]), exp)
    dq.close()
    for sp in streams:
        vx.lib.vx_stream_destroy(C.c_void_p(sp))


@pytest.mark.parametrize("n_rows", [2048, 1 << 14])
def test_accumulate_numerators_misaligned_columns(vx, n_rows):
    """Columns whose device pointers are 4- but not 16-byte aligned (read word by
    word), next to aligned ones, in the tiled and the per-thread forms."""
    rng = np.random.default_rng(n_rows)
    cols = [rnd(rng, n_rows) for _ in range(5)]
    bufs = [vx.DeviceBuffer.from_host(np.concatenate([np.zeros(i % 2, np.uint32), c])) for i, c in enumerate(cols)]
    addrs = np.array([b.addr + 4 * (i % 2) for i, b in enumerate(bufs)], dtype=np.uint64)
    dp = vx.DeviceBuffer(2 * len(addrs))
    vx.lib.vx_memcpy_h2d(dp.ptr, addrs.ctypes.data, 8 * len(addrs))
    idx = np.array([4, 1, 0, 3, 2, 1], np.uint32)
    b, c = rnd(rng, 4 * len(idx)), rnd(rng, 4 * len(idx))
    di, db, dc = vx.DeviceBuffer.from_host(idx), vx.DeviceBuffer.from_host(b), vx.DeviceBuffer.from_host(c)
    outs = [vx.DeviceBuffer(n_rows) for _ in range(4)]
    vx.lib.cuda_accumulate_numerators(dp.ptr, di.ptr, db.ptr, dc.ptr, len(idx), n_rows, *[o.ptr for o in outs])
    vx.synchronize()
    vx.check("cuda_accumulate_numerators")
    assert np.array_equal(np.stack([o.to_host() for o in outs]), O.accumulate_numerators(cols, idx, b, c, n_rows))
