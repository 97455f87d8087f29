"""Blake2s proof-of-work search on the MI355X against the oracle
(cuda/grind.cu, cuda/grind_m31_output.cu, ffi.rs:627-639, 965-976,
grind_ops.rs:20-180, channel.rs:117-168), bit-exact: every search returns the
smallest valid nonce of its range, which is what the oracle computes."""

import ctypes as C
import struct

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu
P = (1 << 31) - 1
NONE = (1 << 64) - 1


def rand_state(seed):
    return np.random.default_rng(seed).integers(0, 1 << 32, 8, dtype=np.uint64).astype(np.uint32)


def u64_buf(vx, value=NONE):
    b = vx.DeviceBuffer(2)
    h = np.array([value], np.uint64)
    vx.lib.vx_memcpy_h2d(b.ptr, h.ctypes.data, 8)
    return b


def read_u64(b) -> int:
    lo, hi = b.to_host(2)
    return int(lo) | (int(hi) << 32)


def abi_search(vx, prefixed, bits, m31, offset, n, init=NONE):
    fn = vx.lib.cuda_grind_pow_m31_output if m31 else vx.lib.cuda_grind_pow
    d_pre, res = vx.DeviceBuffer.from_host(prefixed), u64_buf(vx, init)
    fn(d_pre.ptr, res.ptr, bits, offset, n)
    vx.synchronize()
    vx.check("grind")
    return read_u64(res)


@pytest.mark.parametrize("m31", [False, True])
@pytest.mark.parametrize("bits,offset,n", [(0, 0, 1000), (0, 77, 1), (8, 0, 1 << 16), (12, 5, 100003),
                                           (16, (1 << 32) - 40000, 1 << 20), (20, 123456789, 1 << 22)])
def test_abi_batch_matches_oracle(vx, m31, bits, offset, n):
    pre = O.pow_prefix(rand_state(bits + 3 * m31), bits, m31)
    exp = O.pow_search(pre, bits, m31, offset, n)
    print(f"grind abi m31={m31} bits={bits} offset={offset} n={n}", flush=True)
    assert abi_search(vx, pre, bits, m31, offset, n) == exp


@pytest.mark.parametrize("m31", [False, True])
def test_reference_batch_loop(vx, m31):
    """The reference's host loop (grind_ops.rs:70-115; prover.rs:2474-2496):
    batches of 2^22 nonces into one result until one is found."""
    bits = 22
    pre = O.pow_prefix(rand_state(50 + m31), bits, m31)
    exp = O.pow_search(pre, bits, m31, 0, 1 << 28)
    assert exp != NONE
    fn = vx.lib.cuda_grind_pow_m31_output if m31 else vx.lib.cuda_grind_pow
    d_pre, res = vx.DeviceBuffer.from_host(pre), u64_buf(vx)
    off, batches = 0, 0
    while True:
        fn(d_pre.ptr, res.ptr, bits, off, 1 << 22)
        vx.synchronize()
        batches += 1
        got = read_u64(res)
        if got != NONE:
            break
        off += 1 << 22
        assert batches < 64
    assert got == exp and batches == exp // (1 << 22) + 1
    # a later batch leaves the result alone
    fn(d_pre.ptr, res.ptr, bits, off + (1 << 22), 1 << 22)
    vx.synchronize()
    assert read_u64(res) == exp
    vx.check()


def test_existing_result_bounds_the_search(vx):
    bits = 10
    pre = O.pow_prefix(rand_state(60), bits)
    first = O.pow_search(pre, bits, False, 0, 1 << 20)
    second = O.pow_search(pre, bits, False, first + 1, 1 << 20)
    # a result already below the range: untouched
    assert abi_search(vx, pre, bits, False, second, 1 << 20, init=first) == first
    # a result inside the range: lowered to the first valid nonce
    assert abi_search(vx, pre, bits, False, 0, 1 << 20, init=second) == first
    # an impossible target finds nothing
    assert abi_search(vx, pre, 129, False, 0, 1 << 20) == NONE


@pytest.mark.parametrize("m31", [False, True])
def test_wide_targets_by_chaining_narrow_searches(vx, m31):
    """pow_bits > 32 runs the four-word check, out of the oracle's reach at
    useful sizes. Every nonce with >= 33 zeros has >= 32, so the first 33-bit
    nonce is the first of the 32-bit ones (found one after another by the
    word-0 kernel) whose hashlib digest has >= 33 zeros."""
    import hashlib

    pre = O.pow_prefix(rand_state(70 + m31), 33, m31)

    def zeros(nonce):
        w = np.frombuffer(hashlib.blake2s(pre.tobytes() + struct.pack("<Q", nonce)).digest()[:16], np.uint32)
        if m31:
            w = w % P
        x = int.from_bytes(w.astype("<u4").tobytes(), "little")
        return 128 if x == 0 else (x & -x).bit_length() - 1

    span = 1 << 37
    start, first33 = 0, None
    for _ in range(32):
        n32 = abi_search(vx, pre, 32, m31, start, min(span - start, (1 << 32) - 1))
        print(f"grind wide m31={m31}: 32-bit nonce {n32}", flush=True)
        if n32 == NONE:
            start += (1 << 32) - 1
            if start >= span:
                break
            continue
        assert zeros(n32) >= 32
        if zeros(n32) >= 33:
            first33 = n32
            break
        start = n32 + 1
    assert first33 is not None, "no 33-bit nonce below 2^37"
    # the four-word kernel finds the same nonce in one search
    d_pre, res = vx.DeviceBuffer.from_host(pre), u64_buf(vx)
    fn = vx.lib.cuda_grind_pow_m31_output if m31 else vx.lib.cuda_grind_pow
    off = 0
    while off <= first33:
        fn(d_pre.ptr, res.ptr, 33, off, (1 << 32) - 1)
        off += (1 << 32) - 1
    vx.synchronize()
    assert read_u64(res) == first33
    vx.check()


def dev_channel(vx, state, n_draws=5):
    words = np.zeros(16, np.uint32)
    words[:8] = state
    words[8] = n_draws
    return vx.DeviceBuffer.from_host(words)


@pytest.mark.parametrize("m31", [False, True])
@pytest.mark.parametrize("bits,start,count", [(0, 9, 10), (14, 0, NONE), (18, 1 << 33, 1 << 40), (21, 3, 1 << 30)])
def test_device_grind_from_channel_state(vx, m31, bits, start, count):
    st = rand_state(80 + bits)
    pre = O.pow_prefix(st, bits, m31)
    exp = O.pow_search(pre, bits, m31, start, min(count, 1 << 28))
    assert exp != NONE
    ch, res = dev_channel(vx, st), u64_buf(vx, 0)  # result is reset by the call
    vx.grind_device(ch.addr, bits, res.addr, m31_output=m31, start=start, count=count)
    vx.synchronize()
    assert read_u64(res) == exp
    # the transcript is untouched without mix
    assert np.array_equal(ch.to_host()[:9], np.concatenate([st, [5]]).astype(np.uint32))


def test_device_grind_mixes_nonce_into_transcript(vx):
    st, bits = rand_state(90), 16
    ch, res = dev_channel(vx, st), u64_buf(vx)
    vx.grind_device(ch.addr, bits, res.addr, mix=True)
    vx.synchronize()
    nonce = read_u64(res)
    assert nonce == O.pow_search(O.pow_prefix(st, bits), bits, False, 0, 1 << 26)
    oc = O.ChannelStruct()
    C.memmove(oc.state, st.tobytes(), 32)
    O.lib().orc_channel_mix_u64(C.byref(oc), nonce)
    got = ch.to_host()
    assert got[:8].tobytes() == bytes(oc.state) and got[8] == 0
    # nothing found in the range (the oracle agrees): the transcript stays as it was
    ch2, res2 = dev_channel(vx, st), u64_buf(vx)
    assert O.pow_search(O.pow_prefix(st, 40), 40, False, 0, 1 << 16) == NONE
    vx.grind_device(ch2.addr, 40, res2.addr, count=1 << 16, mix=True)
    vx.synchronize()
    assert read_u64(res2) == NONE
    assert np.array_equal(ch2.to_host()[:9], np.concatenate([st, [5]]).astype(np.uint32))
    vx.check()
    # a target no 128-bit value can meet is refused
    with pytest.raises(vx.HipError):
        vx.grind_device(ch2.addr, 129, res2.addr, count=1 << 16)


def test_python_grind_matches_oracle(vx):
    st = rand_state(100)
    for bits, m31 in [(12, False), (19, True), (24, False)]:
        assert vx.grind(st, bits, m31) == O.pow_search(O.pow_prefix(st, bits, m31), bits, m31, 0, 1 << 30)
    assert vx.grind(st, 64, count=1 << 20) is None
