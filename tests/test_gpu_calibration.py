"""Device calibrations the bench reports beside its rooflines: the int32 VALU
issue rate and the latency of one Merkle node hash on a single wave. Their
values depend on the device; the tests check that they are measured and
plausible. Runs on the MI355X."""

import math

import pytest

pytestmark = pytest.mark.gpu


def test_valu_peak_is_measured(vx):
    tops = vx.measure_valu_peak()
    assert math.isfinite(tops) and 1.0 < tops < 1000.0


def test_hash_latency_is_measured(vx):
    quad, lane = vx.measure_hash_latency()
    # a compression is hundreds of dependent instructions: well above 50 ns and
    # far below a millisecond on any device this library runs on
    for ns in (quad, lane):
        assert math.isfinite(ns) and 50.0 < ns < 1e6
