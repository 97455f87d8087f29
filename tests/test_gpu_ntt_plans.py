"""Every pass plan of the NTT (contiguous depth, strided pass count and row
segment) gives the oracle's results bit for bit, native and stwo basis; so
does every plan of a blowup-4 evaluation, with or without a replicating top
pass, in proofs (single and batched) and wide commitments; the autotuner only
ever installs such plans. Runs on the MI355X."""

import numpy as np
import pytest

import inputs
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
        lde = vx.ntt_lde_plan(k)
        assert sum(m for m, _, _ in lde) == k and lde[0][1] == 0
        assert all(r == 0 for _, _, r in lde[:-1])
    # a proof after tuning is still the oracle's
    assert vx.prove(3, 4, 16, 80, 3) == O.prove(3, 4, 16, 80, 3)


# plans of blowup-4 evaluations, k = log_n + 2: (layers, segment_log, replicating);
# a replicating top pass holds layers - 2 rows of the first quarter in LDS
LDE_PLANS = {
    13: [[(13, 0, 1)], [(8, 0, 0), (5, 4, 1)], [(6, 0, 0), (7, 5, 1)], [(6, 0, 0), (7, 6, 1)],
         [(4, 0, 0), (4, 4, 0), (5, 4, 1)], [(8, 0, 0), (5, 5, 0)]],
    16: [[(16, 0, 1)], [(12, 0, 0), (4, 5, 1)], [(8, 0, 0), (8, 6, 1)], [(6, 0, 0), (10, 4, 1)],
         [(14, 0, 0), (2, 4, 0)]],
    20: [[(10, 0, 0), (10, 6, 1)], [(9, 0, 0), (11, 5, 1)], [(8, 0, 0), (12, 4, 1)], [(14, 0, 0), (6, 4, 1)],
         [(8, 0, 0), (5, 5, 0), (7, 5, 1)], [(12, 0, 0), (8, 5, 0)]],
}


def test_lde_plan_set_and_reject(vx, reset_plans):
    general = vx.ntt_plan(22)
    vx.ntt_lde_plan_set(22, [(12, 0, 0), (10, 5, 1)])
    assert vx.ntt_lde_plan(22) == [(12, 0, 0), (10, 5, 1)]
    assert vx.ntt_plan(22) == general  # the general table is untouched
    for k, plan in [(22, [(12, 0, 0), (10, 5, 0)]),   # 10 rows at c = 5
                    (22, [(12, 0, 1), (10, 5, 0)]),   # replicating pass below the top
                    (17, [(17, 0, 1)]),               # 2^15 rows
                    (16, [(14, 0, 0), (2, 5, 1)])]:   # no rows left
        with pytest.raises(vx.HipError):
            vx.ntt_lde_plan_set(k, plan)
    with pytest.raises(vx.HipError):
        vx.ntt_plan_set(22, [(12, 0), (10, 5)])  # replication is for LDE plans only
    vx.ntt_lde_plan_set(22, [])
    assert vx.ntt_lde_plan(22) == [(k, c, 0) for k, c in vx.ntt_plan(22)]


@pytest.mark.parametrize("k", sorted(LDE_PLANS))
def test_every_lde_plan_is_bit_exact(vx, reset_plans, k):
    log_n = k - 2
    exp_proof = O.prove(5, 7, log_n, 80, 3)
    exp_lde = O.lde(O.fibonacci_trace(5, 7, log_n), log_n)
    pairs = [(5, 7), (11, 13), (17, 19)]
    exp_batch = [O.prove(a, b, log_n, 80, 3) for a, b in pairs[1:]] if k <= 16 else None
    for plan in LDE_PLANS[k]:
        vx.ntt_lde_plan_set(k, plan)
        assert vx.ntt_lde_plan(k) == plan
        p = vx.Prover(log_n, 80, 3)
        assert p.prove(5, 7) == exp_proof, plan
        assert np.array_equal(p.eval_words(), exp_lde), plan
        p.close()
        if exp_batch is not None:  # grid.y = proof
            bp = vx.Prover(log_n, 80, 3, capacity=4)
            got = bp.prove_batch(pairs)
            assert got[0] == exp_proof and got[1:] == exp_batch, plan
            bp.close()


def test_wide_commit_with_replicating_lde_plans(vx, reset_plans):
    """stwo basis: the wide commitment's LDE of every size through replicating plans"""
    logs, blowup, last = [8, 8, 11, 14], 2, 3
    cols = inputs.m31_columns(31, logs)
    root, fri, lastv, beta = O.wide_commit(cols, logs, blowup, last)
    for plans in ({10: [(10, 0, 1)], 13: [(13, 0, 1)], 16: [(16, 0, 1)]},
                  {10: [(4, 0, 0), (6, 4, 1)], 13: [(6, 0, 0), (7, 6, 1)], 16: [(8, 0, 0), (8, 5, 1)]}):
        for k, plan in plans.items():
            vx.ntt_lde_plan_set(k, plan)
        w = vx.WideCommit(logs, blowup, last)
        bufs = [vx.DeviceBuffer.from_host(c) for c in cols]
        w.commit_device([b.addr for b in bufs])
        g_root, g_fri, g_last, g_beta = w.results()
        assert np.array_equal(g_root, root) and np.array_equal(g_fri, fri), plans
        assert np.array_equal(g_last, lastv) and np.array_equal(g_beta, beta), plans
        for c, l in enumerate(logs):
            assert np.array_equal(w.lde(c), O.stwo_lde(cols[c], l, blowup)), (plans, c)
        w.close()


# LDE bridge: the evaluation plan's replicating top pass over layers [s, k)
# takes the interpolation's layers [s, k - 2) too; with the general plan of the
# interpolation, its passes below s are used when they end at s, else the
# default split of those layers (trace log, general plan, blowup-4 plan)
BRIDGE_PLANS = [
    (11, [(11, 0)], [(13, 0, 1)]),                                   # one contiguous bridge pass
    (11, [(7, 0), (4, 4)], [(7, 0, 0), (6, 4, 1)]),
    (14, [(9, 0), (5, 5)], [(9, 0, 0), (7, 5, 1)]),
    (14, [(12, 0), (2, 4)], [(8, 0, 0), (8, 6, 1)]),                 # no general prefix ends at 8
    (18, [(10, 0), (4, 4), (4, 4)], [(10, 0, 0), (4, 4, 0), (6, 4, 1)]),
]


@pytest.mark.parametrize("bridge", [True, False])
def test_lde_bridge_is_bit_exact(vx, reset_plans, bridge):
    vx.lib.vx_ntt_set_lde_bridge(1 if bridge else 0)
    try:
        expect = {}
        for kt, gen, lde in BRIDGE_PLANS:
            vx.ntt_plan_set(kt, gen)
            vx.ntt_lde_plan_set(kt + 2, lde)
            if kt not in expect:
                expect[kt] = (O.prove(5, 7, kt, 80, 3), O.lde(O.fibonacci_trace(5, 7, kt), kt),
                              [O.prove(a, b, kt, 80, 3) for a, b in ((11, 13), (17, 19))] if kt <= 14 else None)
            proof, lde_words, batch = expect[kt]
            p = vx.Prover(kt, 80, 3)
            assert p.prove(5, 7) == proof, (kt, gen, lde)
            assert np.array_equal(p.eval_words(), lde_words), (kt, gen, lde)
            p.close()
            if batch is not None:
                bp = vx.Prover(kt, 80, 3, capacity=4)
                got = bp.prove_batch([(5, 7), (11, 13), (17, 19)])
                assert got[0] == proof and got[1:] == batch, (kt, gen, lde)
                bp.close()
    finally:
        vx.lib.vx_ntt_set_lde_bridge(1)


@pytest.mark.parametrize("bridge", [True, False])
def test_wide_commit_with_lde_bridge(vx, reset_plans, bridge):
    """stwo basis through the bridge: contiguous (2^10 LDE) and strided tiles"""
    logs, blowup, last = [8, 8, 11, 14], 2, 3
    cols = inputs.m31_columns(33, logs)
    root, fri, lastv, beta = O.wide_commit(cols, logs, blowup, last)
    vx.lib.vx_ntt_set_lde_bridge(1 if bridge else 0)
    try:
        for kt, gen, lde in [(8, [(8, 0)], [(10, 0, 1)]), (11, [(6, 0), (5, 5)], [(6, 0, 0), (7, 5, 1)]),
                             (14, [(8, 0), (6, 6)], [(8, 0, 0), (8, 6, 1)])]:
            vx.ntt_plan_set(kt, gen)
            vx.ntt_lde_plan_set(kt + 2, lde)
        w = vx.WideCommit(logs, blowup, last)
        bufs = [vx.DeviceBuffer.from_host(c) for c in cols]
        w.commit_device([b.addr for b in bufs])
        g_root, g_fri, g_last, g_beta = w.results()
        assert np.array_equal(g_root, root) and np.array_equal(g_fri, fri)
        assert np.array_equal(g_last, lastv) and np.array_equal(g_beta, beta)
        for c, l in enumerate(logs):
            assert np.array_equal(w.lde(c), O.stwo_lde(cols[c], l, blowup)), c
        w.close()
    finally:
        vx.lib.vx_ntt_set_lde_bridge(1)
