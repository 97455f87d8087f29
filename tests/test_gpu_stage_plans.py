"""Every chunk depth of the Merkle stage kernels gives the same proofs and
roots: proofs made under forced depths (shallowest and deepest everywhere,
and the autotuner's choice) equal those under the default rule and the
oracle's. Runs on the MI355X."""

import hashlib

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu
P = (1 << 31) - 1


@pytest.fixture
def reset_stage_plans(vx):
    vx.lib.vx_merkle_stage_plan_reset()
    yield
    vx.lib.vx_merkle_stage_plan_reset()


def proofs(vx, log_n, ab, cap):
    p = vx.Prover(log_n, 80, 3, capacity=cap)
    try:
        return p.prove_batch(ab)
    finally:
        p.close()


def force_all(vx, logc):
    for l in range(0, 40):
        assert vx.lib.vx_merkle_stage_plan_set(l, logc) == 0


@pytest.mark.parametrize("log_n,cap", [(11, 1), (12, 6), (16, 1), (17, 3)])
def test_forced_depths_keep_proofs(vx, reset_stage_plans, log_n, cap):
    ab = [(i + 3, 7 * i + 1) for i in range(cap)]
    want = proofs(vx, log_n, ab, cap)
    assert want[0] == O.prove(*ab[0], log_n, 80, 3)
    for logc in (0, 5):
        force_all(vx, logc)
        assert proofs(vx, log_n, ab, cap) == want, f"depth {logc}"
    vx.lib.vx_merkle_stage_plan_reset()
    assert proofs(vx, log_n, ab, cap) == want


def test_autotune_and_plan_api(vx, reset_stage_plans):
    assert vx.lib.vx_merkle_stage_plan_set(40, 1) != 0 and vx.lib.vx_merkle_stage_plan_set(12, 6) != 0
    assert vx.lib.vx_merkle_autotune(11, 13, None) != 0  # below the smallest staged tree
    vx.lib.vx_get_last_error()
    tuned = vx.merkle_autotune(12, 20)
    for l, (ms, c) in tuned.items():
        assert 0 <= c <= 5 and (1 << l) >> (7 + c) >= 2 and ms > 0
        assert vx.merkle_stage_plan(l) == c
    ab = [(5, 9), (11, 13)]
    assert proofs(vx, 14, ab, 2) == [O.prove(a, b, 14, 80, 3) for a, b in ab]


def test_forced_depths_keep_wide_roots(vx, reset_stage_plans):
    rng = np.random.default_rng(8)
    sizes = [10] * 2 + [12] * 3 + [14] * 4
    cols = [rng.integers(0, P, 1 << l, dtype=np.uint64).astype(np.uint32) for l in sizes]
    bufs = [vx.DeviceBuffer.from_host(c) for c in cols]

    def commit():
        w = vx.WideCommit(sizes, log_blowup=2, log_last_layer=4)
        try:
            w.commit_device([b.addr for b in bufs])
            root, fri_roots, last, beta = w.results()
            return hashlib.sha256(np.concatenate([np.ravel(root), np.ravel(fri_roots), np.ravel(last)])
                                  .astype(np.uint32).tobytes()).hexdigest()
        finally:
            w.close()

    want = commit()
    for logc in (0, 5):
        force_all(vx, logc)
        assert commit() == want, f"depth {logc}"
