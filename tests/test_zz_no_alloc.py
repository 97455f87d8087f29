"""vx_deep_quotient_noalloc, the allocation-free form of the fused DEEP
quotient, through the C ABI: bit-identical to vx_deep_quotient on two calls in
a row on one stream (the second smaller, so stale metadata of the first would
show), and sizes beyond its fixed device scratch are refused on the host."""

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
P = (1 << 31) - 1


def rnd(rng, *shape):
    return rng.integers(0, P, shape, dtype=np.uint64).astype(np.uint32)


def case(vx, rng, L, logs, n_points):
    cols = [vx.DeviceBuffer.from_host(rnd(rng, 1 << l)) for l in logs]
    points = rnd(rng, n_points, 8)
    lc = rnd(rng, n_points, len(logs), 12)
    return cols, np.asarray(logs, np.uint32), points, lc


def noalloc(vx, L, cols, logs, points, lc, dxs, dys, outs, stream=None):
    ptrs = (C.c_void_p * len(cols))(*[c.addr for c in cols])
    return vx.lib.vx_deep_quotient_noalloc(ptrs, logs.ctypes.data, len(cols), points.shape[0], points.ctypes.data,
                                           lc.ctypes.data, L, C.c_void_p(dxs.addr), C.c_void_p(dys.addr),
                                           *[C.c_void_p(o.addr) for o in outs], stream)


def test_noalloc_matches_deep_quotient(vx):
    rng = np.random.default_rng(2027)
    # call 1: 40 columns at 3 points (688 metadata words: two staging chunks)
    L1, L2 = 14, 10
    c1 = case(vx, rng, L1, [14] * 20 + [12] * 12 + [9] * 8, 3)
    c2 = case(vx, rng, L2, [10, 7, 10, 5, 9], 1)
    d1 = vx.circle_domain_points(L1)
    d2 = vx.circle_domain_points(L2)
    out1 = [vx.DeviceBuffer(1 << L1) for _ in range(4)]
    out2 = [vx.DeviceBuffer(1 << L2) for _ in range(4)]
    print("noalloc: two calls in a row, 40 columns 2^14 rows then 5 columns 2^10 rows", flush=True)
    assert noalloc(vx, L1, *c1, *d1, out1) == 0
    assert noalloc(vx, L2, *c2, *d2, out2) == 0
    vx.synchronize()
    vx.check("vx_deep_quotient_noalloc")
    for L, (cols, logs, points, lc), (dxs, dys), got in ((L1, c1, d1, out1), (L2, c2, d2, out2)):
        ref = [vx.DeviceBuffer(1 << L) for _ in range(4)]
        dq = vx.DeepQuotient(len(cols), points.shape[0])
        dq.quotient([c.addr for c in cols], logs, points, lc, L, dxs.addr, dys.addr, [o.addr for o in ref])
        vx.synchronize()
        dq.close()
        assert np.array_equal(np.stack([o.to_host() for o in got]), np.stack([o.to_host() for o in ref])), L
    # the first shape again after the smaller one: the scratch is rewritten, not appended to
    again = [vx.DeviceBuffer(1 << L1) for _ in range(4)]
    assert noalloc(vx, L1, *c1, *d1, again) == 0
    vx.synchronize()
    assert np.array_equal(np.stack([o.to_host() for o in again]), np.stack([o.to_host() for o in out1]))


def test_noalloc_refuses_oversized_metadata(vx):
    rng = np.random.default_rng(5)
    L = 8
    n_cols = 820  # 64 + 20 * 820 words at 4 points > 8192
    col = vx.DeviceBuffer.from_host(rnd(rng, 1 << L))
    cols = [col] * n_cols
    logs = np.full(n_cols, L, np.uint32)
    points = rnd(rng, 4, 8)
    lc = rnd(rng, 4, n_cols, 12)
    dxs, dys = vx.circle_domain_points(L)
    outs = [vx.DeviceBuffer(1 << L) for _ in range(4)]
    vx.lib.vx_get_last_error()
    assert noalloc(vx, L, cols, logs, points, lc, dxs, dys, outs) == -1
    assert vx.lib.vx_get_last_error() == -2
    # 406 columns at 4 points (8184 words) still fit
    assert noalloc(vx, L, cols[:406], logs[:406], points, lc[:, :406], dxs, dys, outs) == 0
    vx.synchronize()
    vx.check("vx_deep_quotient_noalloc")
