"""ctypes access to the oracle (oracle/liboracle.so), the CPU restatement used
as the checker. Test infrastructure only."""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")

u32, u64, i32, vp = C.c_uint32, C.c_uint64, C.c_int32, C.c_void_p
P = (1 << 31) - 1
P32 = C.POINTER(C.c_uint32)


class ChannelStruct(C.Structure):
    _fields_ = [("state", C.c_uint8 * 32), ("n_draws", C.c_uint32)]


_SIGS = {
    "orc_set_threads": (None, [u32]),
    "orc_get_threads": (u32, []),
    "orc_m31_add": (u32, [u32, u32]),
    "orc_m31_sub": (u32, [u32, u32]),
    "orc_m31_mul": (u32, [u32, u32]),
    "orc_m31_inv": (u32, [u32]),
    "orc_cm31_mul": (None, [P32, P32, P32]),
    "orc_qm31_mul": (None, [P32, P32, P32]),
    "orc_qm31_inv": (None, [P32, P32]),
    "orc_coset": (None, [i32, u32, P32]),
    "orc_coset_at": (None, [i32, u32, u64, P32]),
    "orc_circle_mul": (None, [P32, P32, P32]),
    "orc_circle_repeated_double": (None, [P32, u32, P32]),
    "orc_ntt_twiddles": (None, [i32, u32, vp, vp, vp, vp, vp, vp]),
    "orc_ntt_twiddles_lean": (None, [i32, u32, vp, vp, vp, vp]),
    "orc_prove_summary": (C.c_int, [u32, u32, u32, u32, u32, vp, vp, vp]),
    "orc_batch_inverse": (None, [vp, vp, u64]),
    "orc_ntt_layer": (None, [vp, vp, u32, u64, i32]),
    "orc_ntt_evaluate": (None, [vp, vp, vp, u32]),
    "orc_ntt_interpolate": (None, [vp, vp, vp, u32]),
    "orc_lde": (None, [vp, u32, u32, vp]),
    "orc_blake2s": (None, [vp, u64, vp]),
    "orc_blake2s_single": (None, [vp, u64, vp]),
    "orc_blake2s_compress": (None, [P32, P32, u32, u32]),
    "orc_merkle_leaf_hashes": (None, [vp, u32, u64, vp]),
    "orc_merkle_node_layer": (None, [vp, u64, vp]),
    "orc_merkle_root": (None, [vp, u32, u64, P32]),
    "orc_merkle_verify_path": (i32, [P32, P32, u64, P32, u32]),
    "orc_fibonacci_trace": (None, [u32, u32, u32, vp]),
    "orc_fibonacci_quotient": (None, [vp, u64, P32, vp, vp, vp, vp]),
    "orc_fold_twiddles": (None, [i32, u32, i32, vp]),
    "orc_fold_line": (None, [vp, vp, P32, u64, vp]),
    "orc_fold_circle_into_line": (None, [vp, vp, vp, P32, u64]),
    "orc_channel_init": (None, [C.POINTER(ChannelStruct)]),
    "orc_channel_mix_digest": (None, [C.POINTER(ChannelStruct), P32]),
    "orc_channel_mix_felts": (None, [C.POINTER(ChannelStruct), P32, u64]),
    "orc_channel_mix_u64": (None, [C.POINTER(ChannelStruct), u64]),
    "orc_channel_draw_felt": (None, [C.POINTER(ChannelStruct), P32]),
    "orc_channel_draw_number": (u64, [C.POINTER(ChannelStruct), u64]),
    "orc_prove_fibonacci": (u64, [u32, u32, u32, u32, u32, vp, u64, P32]),
    "orc_time_prove_fibonacci": (C.c_double, [u32, u32, u32, u32, u32, u32]),
    "orc_time_prove_each": (C.c_int, [u32, u32, u32, u32, u32, u32, u32, vp]),
    "orc_verify_krak": (i32, [vp, u64, u32, u32]),
    "orc_lifted_row": (u32, [u32, u32, u32]),
    "orc_stwo_twiddles": (None, [u32, vp, vp]),
    "orc_stwo_ntt_evaluate": (None, [vp, vp, u32]),
    "orc_stwo_ntt_interpolate": (None, [vp, vp, u32]),
    "orc_lifted_leaves": (None, [vp, vp, u32, u32, vp]),
    "orc_pack_leaves": (None, [vp, vp, u64]),
    "orc_bit_reverse_qm31": (None, [vp, vp, u64, u32]),
    "orc_wide_commit": (u32, [vp, vp, u32, u32, u32, P32, vp, vp, vp]),
    "orc_time_wide_commit": (C.c_double, [vp, vp, u32, u32, u32]),
    "orc_splitmix64": (u64, [C.POINTER(u64)]),
    "orc_accumulate_numerators": (None, [vp, vp, vp, vp, u32, u32, vp, vp, vp, vp]),
    "orc_quotients_combine": (None, [vp, vp, vp, vp, vp, vp, vp, vp, u32, vp, vp, u32, u32, vp, vp, vp, vp]),
    "orc_accumulate_lift": (None, [vp, vp, u32, u32]),
    "orc_circle_domain_brev": (None, [u32, vp, vp]),
    "orc_pow_reduce_word": (u32, [u32]),
    "orc_pow_prefix": (None, [P32, u32, i32, P32]),
    "orc_pow_zeros": (u32, [P32, u64, i32]),
    "orc_pow_search": (u64, [P32, u32, i32, u64, u64]),
}

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build()
        L = C.CDLL(ORACLE_SO)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def arr(xs) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(xs, dtype=np.uint32))


def p32(a: np.ndarray):
    return a.ctypes.data_as(P32)


def ptr_array(cols):
    """Array of column pointers (keeps the columns alive through the returned tuple)."""
    cols = [arr(c) for c in cols]
    ptrs = (vp * len(cols))(*[c.ctypes.data for c in cols])
    return ptrs, cols


def ntt_twiddles_lean(kind: int, log_n: int):
    """(line, circle, iline, icircle) of the layer-from-layer builder."""
    n = 1 << log_n
    line = np.zeros(max(n - 1, 1), np.uint32)
    circle = np.zeros(max(n // 2, 1), np.uint32)
    iline = np.zeros_like(line)
    icircle = np.zeros_like(circle)
    lib().orc_ntt_twiddles_lean(kind, log_n, line.ctypes.data, circle.ctypes.data, iline.ctypes.data,
                                icircle.ctypes.data)
    return line, circle, iline, icircle


def prove_summary(a: int, b: int, log_n: int, n_queries: int = 80, log_last: int = 3) -> dict:
    """Commitments of prove(a, b, log_n) without trees (bounded host memory):
    trace/quotient roots, FRI roots, last layer (4 words per value), queries."""
    J = log_n + 2 - log_last
    roots = np.zeros(8 * (2 + J), np.uint32)
    last = np.zeros(4 << log_last, np.uint32)
    qs = np.zeros(n_queries, np.uint64)
    rc = lib().orc_prove_summary(a, b, log_n, n_queries, log_last, roots.ctypes.data, last.ctypes.data,
                                 qs.ctypes.data)
    assert rc == 0, "orc_prove_summary rejected its parameters"
    r = roots.reshape(-1, 8)
    return {"trace_root": r[0].tolist(), "quotient_root": r[1].tolist(), "fri_roots": r[2:].tolist(),
            "last_layer": last.reshape(-1, 4).tolist(), "queries": [int(q) for q in qs]}


def ntt_twiddles(kind: int, log_n: int):
    n = 1 << log_n
    line = np.zeros(max(n - 1, 1), np.uint32)
    circle = np.zeros(max(n // 2, 1), np.uint32)
    iline = np.zeros_like(line)
    icircle = np.zeros_like(circle)
    offs = np.zeros(max(log_n - 1, 1), np.uint32)
    sizes = np.zeros_like(offs)
    lib().orc_ntt_twiddles(kind, log_n, line.ctypes.data, circle.ctypes.data, iline.ctypes.data,
                           icircle.ctypes.data, offs.ctypes.data, sizes.ctypes.data)
    return line, circle, iline, icircle, offs, sizes


def lde(trace, log_n: int, log_blowup: int = 2) -> np.ndarray:
    t = arr(trace)
    out = np.zeros(1 << (log_n + log_blowup), np.uint32)
    lib().orc_lde(t.ctypes.data, log_n, log_blowup, out.ctypes.data)
    return out


def fibonacci_trace(a: int, b: int, log_n: int) -> np.ndarray:
    out = np.zeros(1 << log_n, np.uint32)
    lib().orc_fibonacci_trace(a, b, log_n, out.ctypes.data)
    return out


def prove(a: int, b: int, log_n: int, n_queries: int = 80, log_last: int = 3) -> bytes:
    L = lib()
    n = L.orc_prove_fibonacci(a, b, log_n, n_queries, log_last, None, 0, None)
    if n == 0:
        raise ValueError("oracle rejected parameters")
    buf = (C.c_uint8 * n)()
    L.orc_prove_fibonacci(a, b, log_n, n_queries, log_last, buf, n, None)
    return bytes(buf)


def verify(proof: bytes, n_queries: int = 80, log_last: int = 3) -> int:
    buf = (C.c_uint8 * len(proof)).from_buffer_copy(proof)
    return lib().orc_verify_krak(buf, len(proof), n_queries, log_last)


def merkle_root(cols) -> np.ndarray:
    ptrs, keep = ptr_array(cols)
    root = np.zeros(8, np.uint32)
    lib().orc_merkle_root(ptrs, len(keep), keep[0].size, p32(root))
    return root


def leaf_hashes(cols) -> np.ndarray:
    ptrs, keep = ptr_array(cols)
    out = np.zeros(8 * keep[0].size, np.uint32)
    lib().orc_merkle_leaf_hashes(ptrs, len(keep), keep[0].size, out.ctypes.data)
    return out


def quotient(eval_, alpha):
    e = arr(eval_)
    al = arr(alpha)
    qs = [np.zeros(e.size, np.uint32) for _ in range(4)]
    lib().orc_fibonacci_quotient(e.ctypes.data, e.size, p32(al), *[q.ctypes.data for q in qs])
    return np.stack(qs)


def fold_twiddles(kind: int, log_size: int, use_y: bool) -> np.ndarray:
    out = np.zeros(max(1, (1 << log_size) // 2), np.uint32)
    lib().orc_fold_twiddles(kind, log_size, 1 if use_y else 0, out.ctypes.data)
    return out


def fold_line(cols4, tw, alpha) -> np.ndarray:
    src = [arr(c) for c in cols4]
    half = src[0].size // 2
    outs = [np.zeros(half, np.uint32) for _ in range(4)]
    ins = (vp * 4)(*[c.ctypes.data for c in src])
    ous = (vp * 4)(*[o.ctypes.data for o in outs])
    t = arr(tw)
    lib().orc_fold_line(ins, t.ctypes.data, p32(arr(alpha)), half, ous)
    return np.stack(outs)


def fold_circle(cols4, tw, alpha) -> np.ndarray:
    src = [arr(c) for c in cols4]
    half = src[0].size // 2
    dst = [np.zeros(half, np.uint32) for _ in range(4)]
    ins = (vp * 4)(*[c.ctypes.data for c in src])
    ous = (vp * 4)(*[o.ctypes.data for o in dst])
    t = arr(tw)
    lib().orc_fold_circle_into_line(ous, ins, t.ctypes.data, p32(arr(alpha)), half)
    return np.stack(dst)


def splitmix64_stream(seed: int, count: int):
    s = u64(seed)
    return [lib().orc_splitmix64(C.byref(s)) for _ in range(count)]


def stwo_twiddles(log_n: int):
    half = max(1, (1 << log_n) // 2)
    f, i = np.zeros(half, np.uint32), np.zeros(half, np.uint32)
    lib().orc_stwo_twiddles(log_n, f.ctypes.data, i.ctypes.data)
    return f, i


def stwo_lde(col, log_n: int, log_blowup: int) -> np.ndarray:
    """stwo interpolate on the size-2^log_n canonic domain, zero-extend, evaluate on 2^(log_n+blowup)."""
    _, itw = stwo_twiddles(log_n)
    etw, _ = stwo_twiddles(log_n + log_blowup)
    d = np.zeros(1 << (log_n + log_blowup), np.uint32)
    d[: 1 << log_n] = arr(col)
    lib().orc_stwo_ntt_interpolate(d.ctypes.data, itw.ctypes.data, log_n)
    lib().orc_stwo_ntt_evaluate(d.ctypes.data, etw.ctypes.data, log_n + log_blowup)
    return d


def wide_commit(cols, log_sizes, log_blowup: int, log_last: int):
    """(root[8], fri_roots[n][8], last[2^log_last][4], beta[4]) of the batched wide commitment."""
    cols = [arr(c) for c in cols]
    ptrs = (vp * len(cols))(*[c.ctypes.data for c in cols])
    lg = arr(log_sizes)
    root = np.zeros(8, np.uint32)
    S = max(log_sizes) + log_blowup
    fri = np.zeros(8 * S, np.uint32)
    last = np.zeros(4 << log_last, np.uint32)
    beta = np.zeros(4, np.uint32)
    nf = lib().orc_wide_commit(ptrs, lg.ctypes.data, len(cols), log_blowup, log_last, p32(root), fri.ctypes.data,
                               last.ctypes.data, beta.ctypes.data)
    assert nf == S - log_last, "oracle rejected the wide-commit shape"
    return root, fri[: 8 * nf].reshape(-1, 8), last.reshape(-1, 4), beta


def pow_prefix(state, pow_bits: int, m31_output: bool = False) -> np.ndarray:
    """prefixed_digest words of a channel state (8 words) for a PoW of pow_bits."""
    st, out = arr(state), np.zeros(8, np.uint32)
    lib().orc_pow_prefix(p32(st), pow_bits, int(m31_output), p32(out))
    return out


def pow_zeros(prefixed, nonce: int, m31_output: bool = False) -> int:
    return lib().orc_pow_zeros(p32(arr(prefixed)), nonce, int(m31_output))


def pow_search(prefixed, pow_bits: int, m31_output: bool = False, start: int = 0, count: int = 1 << 24) -> int:
    """Smallest valid nonce in [start, start + count), or 2^64 - 1."""
    return lib().orc_pow_search(p32(arr(prefixed)), pow_bits, int(m31_output), start, count)


def accumulate_numerators(cols, col_indices, b, c, n_rows: int) -> np.ndarray:
    """[4, n_rows] SoA partial numerators sum_i (c_i * col[idx_i][row] - b_i)."""
    ptrs, keep = ptr_array(cols)
    idx, bb, cc = arr(col_indices), arr(b), arr(c)
    out = np.zeros((4, n_rows), np.uint32)
    lib().orc_accumulate_numerators(ptrs, idx.ctypes.data, bb.ctypes.data, cc.ctypes.data, len(idx), n_rows,
                                    *[out[k].ctypes.data for k in range(4)])
    return out


def quotients_combine(spx, spy, fla, numers, acc_logs, dxs, dys, lifting_log: int) -> np.ndarray:
    """[4, n_rows] combined quotient; numers[j] is a [4, 2^acc_logs[j]] SoA array."""
    numers = [np.ascontiguousarray(n, dtype=np.uint32) for n in numers]
    pk = [(vp * len(numers))(*[n[k].ctypes.data for n in numers]) for k in range(4)]
    a, b, f, lg = arr(spx), arr(spy), arr(fla), arr(acc_logs)
    dx, dy = arr(dxs), arr(dys)
    out = np.zeros((4, dx.size), np.uint32)
    lib().orc_quotients_combine(a.ctypes.data, b.ctypes.data, f.ctypes.data, pk[0], pk[1], pk[2], pk[3],
                                lg.ctypes.data, len(numers), dx.ctypes.data, dy.ctypes.data, lifting_log, dx.size,
                                *[out[k].ctypes.data for k in range(4)])
    return out


def accumulate_lift(col, curr, log_ratio: int) -> np.ndarray:
    """col (n x 4 QM31, AoS) += curr at the lifted index; returns the new col."""
    c = np.ascontiguousarray(col, dtype=np.uint32).copy()
    lib().orc_accumulate_lift(c.ctypes.data, arr(curr).ctypes.data, c.size // 4, log_ratio)
    return c


def circle_domain_brev(log_n: int):
    xs, ys = np.zeros(1 << log_n, np.uint32), np.zeros(1 << log_n, np.uint32)
    lib().orc_circle_domain_brev(log_n, xs.ctypes.data, ys.ctypes.data)
    return xs, ys
