"""A Poisson stream of CT-like / MR-like rigid registrations through the
registration server, drawn from a size mix with 30% of the requests at
512x512x256 and the rest at 256x256x128 (4 pyramid levels, nhood width 13):
every served result is bit-identical to the same pair registered alone, in
binary32 and in binary16 storage, and the latency percentiles are reported.
The reference registers one pair per process (mirorr.cxx:309-619); it has no
serving loop to compare with."""
import time

import numpy as np
import pytest

from mirorr_amd import capi, phantom
from mirorr_amd.registration import MirorrPyramid
from mirorr_amd.server import RegistrationServer

pytestmark = pytest.mark.gpu


def _params():
    r = MirorrPyramid()
    r.SetTransformType("rigid")
    r.SetNhoodWidth(13)
    r.SetNhoodGap(3)
    r.GetPyramidScheduleTuner().SetMaxLevelNumber(4)
    return r.params


@pytest.fixture(scope="module")
def pool():
    large = [phantom.make_ct_mr_pair(dims=(512, 512, 256), n_ellipsoids=24, seed=31)]
    small = [phantom.make_ct_mr_pair(dims=(256, 256, 128), n_ellipsoids=24, seed=40 + k)
             for k in range(2)]
    return large, small


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_poisson_stream_results_equal_the_pairs_alone(pool, storage):
    large, small = pool
    pairs = large + small
    p = _params()
    # each distinct pair alone (one worker, one request at a time)
    with RegistrationServer(slots=1) as srv:
        alone = [srv.poll(srv.submit(pr.fixed, pr.moving, p, storage=storage)) for pr in pairs]
    for r in alone:
        assert r["status"] == capi.OK, r["error"]
    rng = np.random.default_rng(11)
    n, rate = 16, 12.0  # requests, arrivals per second: several run at once
    picks = [0 if rng.random() < 0.3 else 1 + int(rng.integers(2)) for _ in range(n)]
    picks[n // 2] = 0  # at least one large request
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n))
    arrivals -= arrivals[0]
    with RegistrationServer(slots=4) as srv:
        t0 = time.perf_counter()
        tickets = []
        for a, k in zip(arrivals, picks):
            d = t0 + a - time.perf_counter()
            if d > 0:
                time.sleep(d)
            tickets.append(srv.submit(pairs[k].fixed, pairs[k].moving, p, borrow=True,
                                      storage=storage))
        res = [srv.poll(t) for t in tickets]
        info = srv.info()
    lat = np.array([r["latency"] for r in res]) * 1e3
    print(f"{storage}: {n} requests at {rate}/s ({sum(k == 0 for k in picks)} large): latency "
          f"p50 {np.percentile(lat, 50):.0f} ms p99 {np.percentile(lat, 99):.0f} ms, "
          f"max running {info['max_running']}")
    assert info["max_running"] >= 2
    for r, k in zip(res, picks):
        assert r["status"] == capi.OK, r["error"]
        np.testing.assert_array_equal(r["params"], alone[k]["params"])
        np.testing.assert_array_equal(r["center"], alone[k]["center"])
