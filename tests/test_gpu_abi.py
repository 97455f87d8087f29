"""Parity of every reference-ABI entry point (include/vortex_hip.h, Part 1)
against the oracle, bit-exact, on seeded inputs. Runs on the MI355X."""

import ctypes as C

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu
P = (1 << 31) - 1


def rnd(seed, n, hi=P):
    return np.random.default_rng(seed).integers(0, hi, n, dtype=np.uint64).astype(np.uint32)


def h32(vals):
    return (C.c_uint32 * len(vals))(*[int(v) for v in vals])


def dev_ptr_array(vx, bufs):
    """Device array of device pointers (u32**)."""
    addrs = np.array([b.addr for b in bufs], dtype=np.uint64)
    d = vx.DeviceBuffer(2 * len(bufs))
    vx.lib.vx_memcpy_h2d(d.ptr, addrs.ctypes.data, 8 * len(bufs))
    return d


def test_m31_elementwise(vx):
    a, b = rnd(1, 5000), rnd(2, 5000)
    da, db, do = vx.DeviceBuffer.from_host(a), vx.DeviceBuffer.from_host(b), vx.DeviceBuffer(5000)
    vx.lib.cuda_m31_add(da.ptr, db.ptr, do.ptr, 5000)
    vx.synchronize()
    assert np.array_equal(do.to_host(), ((a.astype(np.uint64) + b) % P).astype(np.uint32))
    vx.lib.cuda_m31_mul(da.ptr, db.ptr, do.ptr, 5000)
    vx.synchronize()
    assert np.array_equal(do.to_host(), ((a.astype(np.uint64) * b) % P).astype(np.uint32))
    vx.check()


@pytest.mark.parametrize("k", [1, 2, 3, 5, 8, 12, 14, 15, 17, 20])
def test_ntt_evaluate_interpolate(vx, k):
    n = 1 << k
    line, circle, iline, icircle, offs, sizes = O.ntt_twiddles(1, k)
    x = rnd(100 + k, n)
    exp_eval = x.copy()
    O.lib().orc_ntt_evaluate(exp_eval.ctypes.data, line.ctypes.data, circle.ctypes.data, k)
    d = vx.DeviceBuffer.from_host(x)
    dl, dc = vx.DeviceBuffer.from_host(line), vx.DeviceBuffer.from_host(circle)
    dil, dic = vx.DeviceBuffer.from_host(iline), vx.DeviceBuffer.from_host(icircle)
    ho, hs = h32(offs[: k - 1]), h32(sizes[: k - 1])
    vx.lib.cuda_circle_ntt_evaluate(d.ptr, dl.ptr, dc.ptr, ho, hs, k - 1, n)
    vx.check("evaluate")
    assert np.array_equal(d.to_host(), exp_eval)
    vx.lib.cuda_circle_ntt_interpolate(d.ptr, dil.ptr, dic.ptr, ho, hs, k - 1, n)
    vx.check("interpolate")
    assert np.array_equal(d.to_host(), x)


def test_ntt_batch_layer_bitreverse(vx):
    k, n, cols = 11, 1 << 11, 5
    line, circle, iline, icircle, offs, sizes = O.ntt_twiddles(1, k)
    xs = [rnd(200 + c, n) for c in range(cols)]
    bufs = [vx.DeviceBuffer.from_host(x) for x in xs]
    dptrs = dev_ptr_array(vx, bufs)
    dl, dc = vx.DeviceBuffer.from_host(line), vx.DeviceBuffer.from_host(circle)
    dil, dic = vx.DeviceBuffer.from_host(iline), vx.DeviceBuffer.from_host(icircle)
    ho, hs = h32(offs), h32(sizes)
    vx.lib.cuda_circle_ntt_evaluate_batch(dptrs.ptr, dl.ptr, dc.ptr, ho, hs, k - 1, n, cols)
    vx.check("evaluate_batch")
    for x, b in zip(xs, bufs):
        e = x.copy()
        O.lib().orc_ntt_evaluate(e.ctypes.data, line.ctypes.data, circle.ctypes.data, k)
        assert np.array_equal(b.to_host(), e)
    vx.lib.cuda_circle_ntt_interpolate_batch(dptrs.ptr, dil.ptr, dic.ptr, ho, hs, k - 1, n, cols)
    vx.check("interpolate_batch")
    for x, b in zip(xs, bufs):
        assert np.array_equal(b.to_host(), x)
    # layer by layer equals the monolithic transform (prover.rs:2451-2534)
    d = vx.DeviceBuffer.from_host(xs[0])
    for layer in range(k - 2, -1, -1):
        tw = vx.DeviceBuffer.from_host(line[offs[layer]: offs[layer] + sizes[layer]])
        vx.lib.cuda_circle_ntt_layer(d.ptr, tw.ptr, layer + 1, n, 1)
    vx.lib.cuda_circle_ntt_layer(d.ptr, dc.ptr, 0, n, 1)
    vx.synchronize()
    e = xs[0].copy()
    O.lib().orc_ntt_evaluate(e.ctypes.data, line.ctypes.data, circle.ctypes.data, k)
    assert np.array_equal(d.to_host(), e)
    # bit reverse
    d = vx.DeviceBuffer.from_host(xs[1])
    vx.lib.cuda_bit_reverse_m31(d.ptr, k)
    vx.synchronize()
    idx = np.array([int(format(i, f"0{k}b")[::-1], 2) for i in range(n)])
    assert np.array_equal(d.to_host(), xs[1][idx])


@pytest.mark.parametrize("k", [3, 10, 17])
def test_twiddle_entry_points(vx, k):
    n = 1 << k
    cos = (C.c_uint32 * 4)()
    O.lib().orc_coset(1, k, cos)
    ix, iy, sx, sy = list(cos)
    dx, dy = vx.DeviceBuffer(n), vx.DeviceBuffer(n)
    vx.lib.cuda_compute_coset_points(ix, iy, sx, sy, dx.ptr, dy.ptr, n)
    vx.synchronize()
    ex = np.array([0] * n, np.uint32)
    out = (C.c_uint32 * 2)()
    for i in (0, 1, n // 3, n - 1):
        O.lib().orc_coset_at(1, k, i, out)
        assert dx.to_host()[i] == out[0] and dy.to_host()[i] == out[1]
    # extract_and_squash chain reproduces the oracle's line twiddles
    line, circle, iline, icircle, offs, sizes = O.ntt_twiddles(1, k)
    assert np.array_equal(dy.to_host()[: n // 2], circle)
    cur = dx
    dline = vx.DeviceBuffer(n - 1)
    for layer in range(k - 1):
        half = (n >> layer) // 2
        sq = vx.DeviceBuffer(half)
        twv = C.c_void_p(dline.addr + 4 * int(offs[layer]))
        vx.lib.cuda_extract_and_squash(cur.ptr, twv, sq.ptr, half)
        cur = sq
    vx.synchronize()
    got = dline.to_host()
    assert np.array_equal(got[: n - 2], line[: n - 2])
    # squash_x
    sq = vx.DeviceBuffer(n // 2)
    vx.lib.cuda_squash_x(dx.ptr, sq.ptr, n // 2)
    vx.synchronize()
    xs = dx.to_host().astype(np.uint64)
    exp = ((2 * xs[0::2] * xs[0::2] + P - 1) % P).astype(np.uint32)
    assert np.array_equal(sq.to_host(), exp)
    # batch inverse (0 -> 0) and fold twiddle sources
    vals = rnd(7, n)
    vals[5] = 0
    dv, di = vx.DeviceBuffer.from_host(vals), vx.DeviceBuffer(n)
    vx.lib.cuda_batch_inverse_m31(dv.ptr, di.ptr, n)
    vx.synchronize()
    e = np.zeros(n, np.uint32)
    O.lib().orc_batch_inverse(vals.ctypes.data, e.ctypes.data, n)
    assert np.array_equal(di.to_host(), e)
    for kind, use_y in ((1, 1), (2, 0)):
        O.lib().orc_coset(kind, k, cos)
        src = vx.DeviceBuffer(n // 2)
        vx.lib.cuda_compute_fold_twiddle_sources(cos[0], cos[1], cos[2], cos[3], src.ptr, n, k, use_y)
        inv = vx.DeviceBuffer(n // 2)
        vx.lib.cuda_batch_inverse_m31(src.ptr, inv.ptr, n // 2)
        vx.synchronize()
        assert np.array_equal(inv.to_host(), O.fold_twiddles(kind, k, bool(use_y)))
    vx.check()


@pytest.mark.parametrize("n", [1, 63, 64, 65, 1000, 64 * 4096 + 3])
def test_batch_inverse_zero_chunks(vx, n):
    """cuda_batch_inverse_m31 against the oracle with zeros at chunk starts,
    inside chunks and filling chunks: a 64-word chunk led by a zero maps to
    zero entirely (batch_inverse.cu:20-24, 40), every other zero to zero."""
    vals = rnd(900 + n, n)
    rng = np.random.default_rng(n)
    vals[rng.integers(0, n, max(1, n // 50))] = 0
    vals[::64][::3] = 0  # every third chunk starts with a zero
    if n > 200:
        vals[64:128] = 0
    dv, di = vx.DeviceBuffer.from_host(vals), vx.DeviceBuffer(n)
    vx.lib.cuda_batch_inverse_m31(dv.ptr, di.ptr, n)
    vx.synchronize()
    e = np.zeros(n, np.uint32)
    O.lib().orc_batch_inverse(vals.ctypes.data, e.ctypes.data, n)
    assert np.array_equal(di.to_host(), e)
    vx.check()


def test_subgroup_fold_twiddles(vx):
    """Circle-fold twiddles on subgroup(k), the fold domain of a 2^31-point
    evaluation: the identity's y = 0 leads chunk 0, which the batch inverse
    clears (the oracle and its pins give the same words)."""
    k = 14
    n = 1 << k
    cos = (C.c_uint32 * 4)()
    O.lib().orc_coset(0, k, cos)
    src = vx.DeviceBuffer(n // 2)
    vx.lib.cuda_compute_fold_twiddle_sources(cos[0], cos[1], cos[2], cos[3], src.ptr, n, k, 1)
    inv = vx.DeviceBuffer(n // 2)
    vx.lib.cuda_batch_inverse_m31(src.ptr, inv.ptr, n // 2)
    vx.synchronize()
    got = inv.to_host()
    assert np.array_equal(got, O.fold_twiddles(0, k, True))
    assert np.all(got[:64] == 0) and np.all(got[64:] != 0)
    vx.check()


@pytest.mark.parametrize("k", [3, 9, 16])
def test_fold_and_quotient_entry_points(vx, k):
    n = 1 << k
    cols = [rnd(300 + c, n) for c in range(4)]
    alpha = list(rnd(310, 4))
    tw = O.fold_twiddles(2, k, False)
    ins = [vx.DeviceBuffer.from_host(c) for c in cols]
    outs = [vx.DeviceBuffer(n // 2) for _ in range(4)]
    dtw = vx.DeviceBuffer.from_host(tw)
    vx.lib.cuda_fold_line_soa(*[b.ptr for b in ins], dtw.ptr, *[b.ptr for b in outs], h32(alpha), n // 2)
    vx.synchronize()
    assert np.array_equal(np.stack([o.to_host() for o in outs]), O.fold_line(cols, tw, alpha))
    twc = O.fold_twiddles(1, k, True)
    dtwc = vx.DeviceBuffer.from_host(twc)
    dst = [vx.DeviceBuffer.from_host(np.zeros(n // 2, np.uint32)) for _ in range(4)]
    a2 = (C.c_uint32 * 4)()
    O.lib().orc_qm31_mul(h32(alpha), h32(alpha), a2)
    vx.lib.cuda_fold_circle_into_line_soa(*[b.ptr for b in dst], *[b.ptr for b in ins], dtwc.ptr, h32(alpha), a2,
                                          n // 2)
    vx.synchronize()
    assert np.array_equal(np.stack([o.to_host() for o in dst]), O.fold_circle(cols, twc, alpha))
    # quotient, chunked quotient, zero pad
    ev = cols[0]
    dev = vx.DeviceBuffer.from_host(ev)
    qs = [vx.DeviceBuffer(n) for _ in range(4)]
    vx.lib.cuda_fibonacci_quotient(dev.ptr, *[q.ptr for q in qs], h32(alpha), n)
    vx.synchronize()
    expq = O.quotient(ev, alpha)
    assert np.array_equal(np.stack([q.to_host() for q in qs]), expq)
    off, cn = n // 4, n // 2
    qc = [vx.DeviceBuffer(cn) for _ in range(4)]
    vx.lib.cuda_fibonacci_quotient_chunk(dev.ptr, *[q.ptr for q in qc], h32(alpha), off, cn, n)
    vx.synchronize()
    assert np.array_equal(np.stack([q.to_host() for q in qc]), expq[:, off: off + cn])
    pad = vx.DeviceBuffer(4 * n)
    vx.lib.cuda_zero_pad(dev.ptr, pad.ptr, n, 4 * n)
    vx.synchronize()
    assert np.array_equal(pad.to_host(), np.concatenate([ev, np.zeros(3 * n, np.uint32)]))
    vx.check()


@pytest.mark.parametrize("n", [1, 2, 3, 5])
def test_quotient_wraps_tiny_domains(vx, n):
    """(i + 1) mod n and (i + 2) mod n for domains shorter than the stencil
    (constraint.cu:10-36 takes full remainders); checked against the oracle."""
    ev = rnd(320 + n, n)
    alpha = list(rnd(330, 4))
    dev = vx.DeviceBuffer.from_host(ev)
    qs = [vx.DeviceBuffer(n) for _ in range(4)]
    vx.lib.cuda_fibonacci_quotient(dev.ptr, *[q.ptr for q in qs], h32(alpha), n)
    vx.synchronize()
    assert np.array_equal(np.stack([q.to_host() for q in qs]), O.quotient(ev, alpha))
    vx.check()


@pytest.mark.parametrize("n_cols,k", [(1, 6), (4, 10), (16, 11), (20, 10), (3, 12)])
def test_merkle_entry_points(vx, n_cols, k):
    n = 1 << k
    cols = [rnd(400 + c, n) for c in range(n_cols)]
    bufs = [vx.DeviceBuffer.from_host(c) for c in cols]
    dptrs = dev_ptr_array(vx, bufs)
    leaves = vx.DeviceBuffer(8 * n)
    vx.lib.cuda_merkle_hash_leaves(dptrs.ptr, leaves.ptr, n_cols, n)
    vx.synchronize()
    exp_leaves = O.leaf_hashes(cols)
    assert np.array_equal(leaves.to_host(), exp_leaves)
    par = vx.DeviceBuffer(4 * n)
    vx.lib.cuda_merkle_hash_nodes(leaves.ptr, par.ptr, n // 2)
    vx.synchronize()
    exp_par = np.zeros(4 * n, np.uint32)
    O.lib().orc_merkle_node_layer(exp_leaves.ctypes.data, n // 2, exp_par.ctypes.data)
    assert np.array_equal(par.to_host(), exp_par)
    root = vx.DeviceBuffer(8)
    vx.lib.cuda_merkle_reduce_to_root(par.ptr, root.ptr, min(n // 2, 1024))
    vx.synchronize()
    exp_root = O.merkle_root(cols) if n // 2 <= 1024 else None
    if exp_root is not None:
        assert np.array_equal(root.to_host(), exp_root)
    if n >= 1024:
        tiles = vx.DeviceBuffer(8 * (n // 1024))
        vx.lib.cuda_merkle_tiled_generic(dptrs.ptr, tiles.ptr, n_cols, n)
        vx.synchronize()
        got = tiles.to_host().reshape(-1, 8)
        for t in range(n // 1024):
            assert np.array_equal(got[t], O.merkle_root([c[1024 * t: 1024 * (t + 1)] for c in cols]))
    vx.check()


@pytest.mark.parametrize("k", [1, 5, 11, 12, 13])
def test_merkle_soa4_entry_points(vx, k):
    n = 1 << k
    cols = [rnd(500 + c, n) for c in range(4)]
    bufs = [vx.DeviceBuffer.from_host(c) for c in cols]
    exp_root = O.merkle_root(cols)
    if n <= 2048:
        root = vx.DeviceBuffer(8)
        vx.lib.cuda_merkle_commit_small_soa4(*[b.ptr for b in bufs], root.ptr, n)
        vx.synchronize()
        assert np.array_equal(root.to_host(), exp_root)
    par = vx.DeviceBuffer(4 * n)
    vx.lib.cuda_merkle_hash_leaves_merge_soa4(*[b.ptr for b in bufs], par.ptr, n)
    vx.synchronize()
    exp_leaves = O.leaf_hashes(cols)
    exp_par = np.zeros(4 * n, np.uint32)
    O.lib().orc_merkle_node_layer(exp_leaves.ctypes.data, n // 2, exp_par.ctypes.data)
    assert np.array_equal(par.to_host(), exp_par)
    if n >= 1024:
        tiles = vx.DeviceBuffer(8 * (n // 1024))
        vx.lib.cuda_merkle_tiled_soa4(*[b.ptr for b in bufs], tiles.ptr, n)
        vx.synchronize()
        got = tiles.to_host().reshape(-1, 8)
        for t in range(n // 1024):
            assert np.array_equal(got[t], O.merkle_root([c[1024 * t: 1024 * (t + 1)] for c in cols]))
    vx.check()
