"""Parity of the stwo-basis entry points and of the batched wide commitment
against the oracle, bit-exact, on seeded inputs. Runs on the MI355X."""

import ctypes as C
import hashlib
import json
import os

import numpy as np
import pytest

import inputs
import oracle_lib as O

pytestmark = pytest.mark.gpu
P = (1 << 31) - 1
HERE = os.path.dirname(os.path.abspath(__file__))


def rnd(seed, n):
    return np.random.default_rng(seed).integers(0, P, n, dtype=np.uint64).astype(np.uint32)


def dev_ptr_array(vx, addrs):
    a = np.array(addrs, dtype=np.uint64)
    d = vx.DeviceBuffer(2 * max(1, len(addrs)))
    if len(addrs):
        vx.lib.vx_memcpy_h2d(d.ptr, a.ctypes.data, 8 * len(addrs))
    return d


def wide_golden(label):
    with open(os.path.join(HERE, "golden", "wide.json")) as f:
        return next(c for c in json.load(f) if c["label"] == label)


@pytest.mark.parametrize("k", [1, 2, 3, 8, 13, 16, 22])
def test_stwo_twiddles_on_device(vx, k):
    f, i = O.stwo_twiddles(k)
    df, di = vx.stwo_twiddles_device(k)
    assert np.array_equal(df.to_host(), f)
    assert np.array_equal(di.to_host(), i)
    vx.check()


@pytest.mark.parametrize("k", [2, 3, 4, 7, 11, 12, 13, 16, 19, 20, 22])
def test_stwo_ntt_entry_points(vx, k):
    n = 1 << k
    f, i = O.stwo_twiddles(k)
    x = rnd(700 + k, n)
    exp = x.copy()
    O.lib().orc_stwo_ntt_evaluate(exp.ctypes.data, f.ctypes.data, k)
    d = vx.DeviceBuffer.from_host(x)
    df, di = vx.DeviceBuffer.from_host(f), vx.DeviceBuffer.from_host(i)
    vx.lib.cuda_stwo_ntt_evaluate(d.ptr, df.ptr, n)
    vx.check("cuda_stwo_ntt_evaluate")
    assert np.array_equal(d.to_host(), exp)
    vx.lib.cuda_stwo_ntt_interpolate(d.ptr, di.ptr, n)
    vx.check("cuda_stwo_ntt_interpolate")
    assert np.array_equal(d.to_host(), x)


def test_stwo_ntt_suffix_of_larger_buffer(vx):
    # a smaller domain's buffer is the tail of a larger one (stwo slices twiddles that way)
    big, _ = O.stwo_twiddles(16)
    small, _ = O.stwo_twiddles(9)
    assert np.array_equal(big[-len(small):], small)
    x = rnd(3, 1 << 9)
    exp = x.copy()
    O.lib().orc_stwo_ntt_evaluate(exp.ctypes.data, small.ctypes.data, 9)
    d, db = vx.DeviceBuffer.from_host(x), vx.DeviceBuffer.from_host(big)
    vx.lib.cuda_stwo_ntt_evaluate(d.ptr, C.c_void_p(db.addr + 4 * (len(big) - len(small))), 1 << 9)
    vx.check()
    assert np.array_equal(d.to_host(), exp)


def test_stwo_ntt_rejects_tiny_sizes(vx):
    d = vx.DeviceBuffer(2)
    vx.lib.cuda_stwo_ntt_evaluate(d.ptr, d.ptr, 2)
    assert vx.lib.vx_get_last_error() != 0
    vx.lib.cuda_stwo_ntt_interpolate(d.ptr, d.ptr, 3)
    assert vx.lib.vx_get_last_error() != 0


def schedule_bytes(log_sizes):
    """LeafHashChunk array (merkle_ops.rs:77-98): columns grouped by ascending size, chunks of 16."""
    words, i = [], 0
    while i < len(log_sizes):
        j = i
        while j < len(log_sizes) and log_sizes[j] == log_sizes[i]:
            j += 1
        for s in range(i, j, 16):
            e = min(s + 16, j)
            idx = list(range(s, e)) + [0] * (16 - (e - s))
            words += idx + [e - s, log_sizes[i]]
        i = j
    return np.array(words, np.uint32), len(words) // 18


@pytest.mark.parametrize("shape", ["reference", "single", "seventeen", "empty"])
def test_build_leaves_lifted(vx, shape):
    if shape == "reference":  # tests.rs:208-254 mixed shape, scaled down
        logs = [7] * 5 + [8] * 9 + [10] * 17 + [12] * 7
    elif shape == "single":
        logs = [11]
    elif shape == "seventeen":  # 17 equal columns: a full block then a 1-word block
        logs = [9] * 17
    else:
        logs = []
    S = max(logs) if logs else 10
    cols = [rnd(900 + c, 1 << l) for c, l in enumerate(logs)]
    bufs = [vx.DeviceBuffer.from_host(c) for c in cols]
    dptrs = dev_ptr_array(vx, [b.addr for b in bufs])
    sched, n_chunks = schedule_bytes(logs)
    dsched = vx.DeviceBuffer.from_host(sched if len(sched) else np.zeros(1, np.uint32))
    out = vx.DeviceBuffer(8 << S)
    vx.lib.cuda_build_leaves_lifted(dptrs.ptr, dsched.ptr, n_chunks, S, out.ptr, 1 << S)
    vx.synchronize()
    vx.check()
    exp = np.zeros(8 << S, np.uint32)
    if cols:
        ptrs = (C.c_void_p * len(cols))(*[c.ctypes.data for c in cols])
        lg = np.array(logs, np.uint32)
        O.lib().orc_lifted_leaves(ptrs, lg.ctypes.data, len(cols), S, exp.ctypes.data)
    else:
        e = np.frombuffer(hashlib.blake2s(b"").digest(), np.uint32)
        exp = np.tile(e, 1 << S)
    assert np.array_equal(out.to_host(), exp)


@pytest.mark.parametrize("n", [16, 4096, 1000])
def test_pack_leaves_and_bit_reverse_qm31(vx, n):
    ins = [rnd(50 + c, n) for c in range(4)]
    rows = (n + 15) // 16
    outs_exp = [np.zeros(rows, np.uint32) for _ in range(64)]
    ip = (C.c_void_p * 4)(*[a.ctypes.data for a in ins])
    op = (C.c_void_p * 64)(*[a.ctypes.data for a in outs_exp])
    O.lib().orc_pack_leaves(ip, op, n)
    din = [vx.DeviceBuffer.from_host(a) for a in ins]
    dout = [vx.DeviceBuffer(rows) for _ in range(64)]
    for b in dout:
        vx.lib.vx_memset(b.ptr, 0, 4 * rows)
    pin, pout = dev_ptr_array(vx, [b.addr for b in din]), dev_ptr_array(vx, [b.addr for b in dout])
    vx.lib.cuda_pack_leaves(pin.ptr, pout.ptr, n)
    vx.synchronize()
    vx.check()
    for c in range(64):
        assert np.array_equal(dout[c].to_host(), outs_exp[c]), c
    if n & (n - 1) == 0:
        k = n.bit_length() - 1
        q = rnd(77, 4 * n)
        exp = np.zeros_like(q)
        O.lib().orc_bit_reverse_qm31(q.ctypes.data, exp.ctypes.data, n, k)
        dq, dr = vx.DeviceBuffer.from_host(q), vx.DeviceBuffer(4 * n)
        vx.lib.cuda_bit_reverse_qm31(dq.ptr, dr.ptr, n, k)
        vx.synchronize()
        vx.check()
        assert np.array_equal(dr.to_host(), exp)


def run_wide(vx, cols, logs, blowup, last):
    w = vx.WideCommit(logs, blowup, last)
    bufs = [vx.DeviceBuffer.from_host(c) for c in cols]
    w.commit_device([b.addr for b in bufs])
    res = w.results()
    return w, bufs, res


@pytest.mark.parametrize("logs,blowup,last", [
    ([3, 3, 4, 5, 5], 2, 2),
    ([4] * 3 + [6] * 17 + [9] * 2, 2, 3),
    ([10] * 16, 1, 4),
    ([2, 12], 3, 5),
    ([8, 9, 10, 11, 12, 13, 14], 2, 4),
])
def test_wide_commit_matches_oracle(vx, logs, blowup, last):
    cols = inputs.m31_columns(sum(logs) + blowup, logs)
    root, fri, lastv, beta = O.wide_commit(cols, logs, blowup, last)
    w, bufs, (g_root, g_fri, g_last, g_beta) = run_wide(vx, cols, logs, blowup, last)
    assert np.array_equal(g_beta, beta)
    assert np.array_equal(g_root, root)
    assert np.array_equal(g_fri, fri)
    assert np.array_equal(g_last, lastv)
    for c, l in enumerate(logs):
        assert np.array_equal(w.lde(c), O.stwo_lde(cols[c], l, blowup)), c
    # a second commitment on the same context (reused workspace, same pointers) repeats exactly
    w.commit_device([b.addr for b in bufs])
    r2 = w.results()
    assert np.array_equal(r2[0], root) and np.array_equal(r2[1], fri)
    w.close()


def test_wide_commit_config_golden(vx):
    g = wide_golden("wide_config")
    logs = g["log_sizes"]
    cols = inputs.m31_columns(g["seed"], logs)
    w, bufs, (root, fri, last, beta) = run_wide(vx, cols, logs, g["log_blowup"], g["log_last"])
    assert root.tolist() == g["root"]
    assert beta.tolist() == g["beta"]
    assert fri.tolist() == g["fri_roots"]
    assert last.tolist() == g["last_layer"]
    for c in range(len(logs)):
        assert hashlib.sha256(w.lde(c).tobytes()).hexdigest() == g["lde_sha256"][c], c
    # new inputs through the same context: pointers change, results follow the data
    cols2 = [c[::-1].copy() for c in cols[:]]
    bufs2 = [vx.DeviceBuffer.from_host(c) for c in cols2]
    w.commit_device([b.addr for b in bufs2])
    r2 = w.results()
    assert r2[0].tolist() != g["root"]
    w.commit_device([b.addr for b in bufs])
    assert w.results()[0].tolist() == g["root"]
    w.close()


def test_wide_commit_rejects_bad_shapes(vx):
    with pytest.raises(vx.HipError):
        vx.WideCommit([5, 4], 2, 2)
    with pytest.raises(vx.HipError):
        vx.WideCommit([5], 2, 7)
