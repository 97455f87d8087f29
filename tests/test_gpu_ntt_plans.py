"""Every pass plan of the NTT (contiguous depth, strided pass count and row
segment) gives the oracle's results bit for bit, native and stwo basis; the
autotuner only ever installs such plans. Runs on the MI355X."""

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu
P = (1 << 31) - 1

PLANS = {
    12: [[(12, 0)], [(8, 0), (4, 4)], [(6, 0), (6, 5)], [(4, 0), (4, 4), (4, 4)], [(6, 0), (6, 6)]],
    16: [[(14, 0), (2, 4)], [(9, 0), (7, 5)], [(8, 0), (8, 4)], [(10, 0), (3, 5), (3, 5)], [(8, 0), (8, 6)]],
    20: [[(14, 0), (6, 5)], [(10, 0), (10, 4)], [(11, 0), (9, 5)], [(8, 0), (6, 4), (6, 5)], [(13, 0), (7, 4)],
         [(12, 0), (8, 6)], [(6, 0), (7, 6), (7, 5)]],
    23: [[(14, 0), (9, 5)], [(13, 0), (10, 4)], [(12, 0), (6, 5), (5, 4)], [(9, 0), (7, 5), (7, 5)],
         [(11, 0), (6, 6), (6, 6)]],
}


def rnd(seed, n):
    return np.random.default_rng(seed).integers(0, P, n, dtype=np.uint64).astype(np.uint32)


@pytest.fixture
def reset_plans(vx):
    yield
    vx.lib.vx_ntt_plan_reset()


@pytest.mark.parametrize("k", sorted(PLANS))
def test_every_plan_is_bit_exact(vx, reset_plans, k):
    n = 1 << k
    line, circle, iline, icircle, offs, sizes = O.ntt_twiddles(1, k)
    sf, si = O.stwo_twiddles(k)
    x = rnd(k, n)
    exp = x.copy()
    O.lib().orc_ntt_evaluate(exp.ctypes.data, line.ctypes.data, circle.ctypes.data, k)
    exp_s = x.copy()
    O.lib().orc_stwo_ntt_evaluate(exp_s.ctypes.data, sf.ctypes.data, k)
    dl, dc = vx.DeviceBuffer.from_host(line), vx.DeviceBuffer.from_host(circle)
    dil, dic = vx.DeviceBuffer.from_host(iline), vx.DeviceBuffer.from_host(icircle)
    dsf, dsi = vx.DeviceBuffer.from_host(sf), vx.DeviceBuffer.from_host(si)
    import ctypes as C
    ho = (C.c_uint32 * (k - 1))(*[int(v) for v in offs[: k - 1]])
    hs = (C.c_uint32 * (k - 1))(*[int(v) for v in sizes[: k - 1]])
    for plan in PLANS[k]:
        vx.ntt_plan_set(k, plan)
        assert vx.ntt_plan(k) == plan
        d = vx.DeviceBuffer.from_host(x)
        vx.lib.cuda_circle_ntt_evaluate(d.ptr, dl.ptr, dc.ptr, ho, hs, k - 1, n)
        vx.check("evaluate")
        assert np.array_equal(d.to_host(), exp), plan
        vx.lib.cuda_circle_ntt_interpolate(d.ptr, dil.ptr, dic.ptr, ho, hs, k - 1, n)
        assert np.array_equal(d.to_host(), x), plan
        vx.lib.cuda_stwo_ntt_evaluate(d.ptr, dsf.ptr, n)
        assert np.array_equal(d.to_host(), exp_s), plan
        vx.lib.cuda_stwo_ntt_interpolate(d.ptr, dsi.ptr, n)
        vx.check("stwo")
        assert np.array_equal(d.to_host(), x), plan


def test_invalid_plans_are_rejected(vx, reset_plans):
    for k, plan in [(16, [(15, 0), (1, 5)]), (16, [(10, 0), (5, 4)]), (16, [(3, 0), (13, 5)]),
                    (16, [(10, 4), (6, 5)]), (16, [(2, 0), (14, 4)])]:
        with pytest.raises(vx.HipError):
            vx.ntt_plan_set(k, plan)
    assert vx.ntt_plan(16) == [(12, 0), (4, 5)]  # default split unchanged


def test_autotune_installs_valid_plans_and_proofs_stay_exact(vx, reset_plans):
    res = vx.ntt_autotune(8, 18)
    for k, (ms, plan) in res.items():
        assert ms > 0 and sum(m for m, _ in plan) == k and plan[0][1] == 0
    # a proof after tuning is still the oracle's
    assert vx.prove(3, 4, 16, 80, 3) == O.prove(3, 4, 16, 80, 3)
