"""The registration server (mirorr_server_* / mirorr_submit / mirorr_poll):
each result is the same alone and in the mix, admission keeps the device
memory within the HBM budget, the admission bound covers what a registration
allocates, and failures come back as results."""
import numpy as np
import pytest

from mirorr_amd import capi, phantom
from mirorr_amd.registration import MirorrPyramid
from mirorr_amd.server import RegistrationServer

pytestmark = pytest.mark.gpu

# dims, spacing, model, nhood width, nhood gap, levels
CASES = [
    ((64, 64, 32), (3.0, 3.0, 3.0), "rigid", 7, 3, 2),
    ((56, 48, 40), (2.5, 3.0, 2.0), "affine", 5, 2, 2),
    ((72, 64, 36), (2.0, 2.0, 3.0), "rigid", 9, 3, 2),
    ((48, 56, 44), (3.0, 2.5, 2.5), "affine", 7, 4, 2),
    ((64, 48, 48), (2.0, 3.0, 2.0), "affine", 9, 3, 3),
    ((40, 40, 40), (3.0, 3.0, 3.0), "rigid", 5, 2, 2),
]


def _pair(i):
    dims, sp, model, _, _, _ = CASES[i]
    return phantom.make_pair(dims=dims, spacing=sp, n_ellipsoids=16, seed=100 + i, model=model)


def _reg(i):
    _, _, model, nw, ng, lev = CASES[i]
    r = MirorrPyramid()
    r.SetTransformType(model)
    r.SetNhoodWidth(nw)
    r.SetNhoodGap(ng)
    r.GetPyramidScheduleTuner().SetMaxLevelNumber(lev)
    return r


def _alone(i, pair):
    r = _reg(i)
    r.SetFixedImage(pair.fixed)
    r.SetMovingImage(pair.moving)
    return r.RunMirorr()


def test_results_alone_and_in_the_mix_are_identical():
    pairs = [_pair(i) for i in range(len(CASES))]
    alone = [_alone(i, p) for i, p in enumerate(pairs)]
    with RegistrationServer(slots=4) as srv:
        print("server: submitting", len(pairs), "requests", flush=True)
        tickets = [srv.submit(p.fixed, p.moving, _reg(i).params, borrow=i % 2 == 1)
                   for i, p in enumerate(pairs)]
        res = [srv.poll(t) for t in tickets]
        info = srv.info()
    for i, (r, (p, c)) in enumerate(zip(res, alone)):
        assert r["status"] == capi.OK, r["error"]
        print(f"request {i}: slot {r['slot']} latency {1e3 * r['latency']:.1f} ms")
        # The pipeline is deterministic (integer atomics only, fixed-order
        # reductions), so a request's result does not depend on what runs
        # beside it.
        np.testing.assert_array_equal(r["params"], p)
        np.testing.assert_array_equal(r["center"], c)
    # The requests ran at the same time: admitted together, and at least two
    # of them overlapped in time.
    assert info["max_running"] >= 2 and info["completed"] == len(pairs)
    spans = sorted((r["t_start"], r["t_end"]) for r in res)
    assert any(spans[k + 1][0] < spans[k][1] for k in range(len(spans) - 1))
    assert len({r["slot"] for r in res}) >= 2


def test_budget_admits_one_request_at_a_time():
    pairs = [_pair(0), _pair(0), _pair(0)]
    with RegistrationServer(slots=1) as srv:
        r0 = srv.poll(srv.submit(pairs[0].fixed, pairs[0].moving, _reg(0).params))
    assert r0["status"] == capi.OK, r0["error"]
    bound = r0["bytes_bound"]
    budget = int(1.5 * bound)  # room for one request, not two
    with RegistrationServer(slots=4, hbm_budget=budget) as srv:
        ts = [srv.submit(p.fixed, p.moving, _reg(0).params) for p in pairs]
        res = [srv.poll(t) for t in ts]
        info = srv.info()
    print("budget", budget, "reserved", info["reserve_peak"], "max running",
          info["max_running"])
    assert info["max_running"] == 1
    # admission kept what the slots may hold within the budget
    assert bound <= info["reserve_peak"] <= budget
    spans = sorted((r["t_start"], r["t_end"]) for r in res)
    assert all(spans[k + 1][0] >= spans[k][1] for k in range(len(spans) - 1))
    for r in res:
        assert r["status"] == capi.OK, r["error"]
        np.testing.assert_array_equal(r["params"], r0["params"])


@pytest.mark.parametrize("variant", ["symmetric", "classic", "masked", "nw13", "fixed_grid"])
def test_memory_bound_covers_the_registration(variant):
    """The admission bound holds what one registration charges on a fresh
    worker, and is not loose by more than 2x."""
    i = 4 if variant != "nw13" else 2
    pair = _pair(i)
    reg = _reg(i)
    p = reg.params
    fm = mm = None
    if variant == "classic":
        p.symmetric = 0
    if variant == "masked":
        fm = (pair.fixed.data > 100).astype(np.uint8) * 255
        mm = np.full(pair.moving.data.shape, 255, np.uint8)
    if variant == "nw13":
        p.nhood_width = 13
    if variant == "fixed_grid":
        p.resampling = capi.RESAMPLE_FIXED
    with RegistrationServer(slots=1) as srv:
        r = srv.poll(srv.submit(pair.fixed, pair.moving, p, fixed_mask=fm, moving_mask=mm))
        info = srv.info()
    assert r["status"] == capi.OK, r["error"]
    print(f"{variant}: bound {r['bytes_bound']} peak {info['hbm_peak']}")
    assert info["hbm_peak"] <= r["bytes_bound"] <= 2 * info["hbm_peak"]


def test_failures_come_back_as_results():
    pair = _pair(5)
    with RegistrationServer(slots=2, hbm_budget=1 << 20) as srv:
        r = srv.poll(srv.submit(pair.fixed, pair.moving, _reg(5).params))
        assert r["status"] == capi.ERR_NOMEM and r["slot"] == -1
        assert "budget" in r["error"]
    with RegistrationServer(slots=2) as srv:
        bad = _reg(5).params
        bad.resampling = 9
        t_bad = srv.submit(pair.fixed, pair.moving, bad)
        t_ok = srv.submit(pair.fixed, pair.moving, _reg(5).params)
        r_bad, r_ok = srv.poll(t_bad), srv.poll(t_ok)
        assert r_bad["status"] == capi.ERR_ARG and "resampling" in r_bad["error"]
        assert r_ok["status"] == capi.OK, r_ok["error"]
        # a ticket is handed out once
        with pytest.raises(capi.MirorrError):
            srv.poll(t_ok)
        with pytest.raises(capi.MirorrError):
            srv.poll(12345)


def test_small_request_overtakes_one_waiting_for_memory():
    """Budget for one large and one small request: the second large request
    waits for memory while the small one, submitted after it, is admitted
    beside the first; every result is still the request's own."""
    big, small = _pair(2), _pair(5)
    with RegistrationServer(slots=1) as srv:
        rb = srv.poll(srv.submit(big.fixed, big.moving, _reg(2).params))
        rs = srv.poll(srv.submit(small.fixed, small.moving, _reg(5).params))
    assert rb["status"] == capi.OK and rs["status"] == capi.OK
    b_big, b_small = rb["bytes_bound"], rs["bytes_bound"]
    assert b_small < b_big
    budget = b_big + b_small + (b_big - b_small) // 2  # one large and one small, not two large
    with RegistrationServer(slots=4, hbm_budget=budget) as srv:
        t1 = srv.submit(big.fixed, big.moving, _reg(2).params)
        t2 = srv.submit(big.fixed, big.moving, _reg(2).params)
        t3 = srv.submit(small.fixed, small.moving, _reg(5).params)
        r1, r2, r3 = srv.poll(t1), srv.poll(t2), srv.poll(t3)
        info = srv.info()
    print(f"budget {budget}: starts {r1['t_start']:.4f} {r2['t_start']:.4f} {r3['t_start']:.4f}, "
          f"reserved {info['reserve_peak']}")
    assert r3["t_start"] < r2["t_start"]  # the small request did not wait behind the large one
    assert r2["t_start"] >= r1["t_end"]   # two large requests never ran together
    assert b_big + b_small <= info["reserve_peak"] <= budget
    for r, ref in ((r1, rb), (r2, rb), (r3, rs)):
        assert r["status"] == capi.OK, r["error"]
        np.testing.assert_array_equal(r["params"], ref["params"])


def test_destroy_completes_the_waiting_tickets():
    """Closing the server while requests wait for a slot completes them with
    MIRORR_ERR_STATE, so threads blocked in poll return; the requests that ran
    keep their results."""
    import threading
    import time
    pair = _pair(2)
    ref = _alone(2, pair)
    srv = RegistrationServer(slots=1)
    tickets = [srv.submit(pair.fixed, pair.moving, _reg(2).params) for _ in range(48)]
    out = {}

    def wait(t):
        out[t] = srv.poll(t)

    threads = [threading.Thread(target=wait, args=(t,)) for t in tickets]
    for th in threads:
        th.start()
    time.sleep(0.3)  # the pollers are blocked inside mirorr_poll
    srv.close()
    for th in threads:
        th.join(timeout=60)
    assert not any(th.is_alive() for th in threads)
    assert sorted(out) == sorted(tickets)
    ok = [r for r in out.values() if r["status"] == capi.OK]
    shut = [r for r in out.values() if r["status"] == capi.ERR_STATE]
    print(f"ran {len(ok)}, shut down {len(shut)}")
    assert len(ok) + len(shut) == len(tickets) and ok and shut
    assert all("shut down" in r["error"] and r["slot"] == -1 for r in shut)
    for r in ok:
        np.testing.assert_array_equal(r["params"], ref[0])


def test_batched_slots_give_each_request_its_alone_result():
    # Requests with equal parameters are batched on a slot (one search launch
    # per pass over every pair's tiles); the others run alone. Three groups:
    # rigid NW 7 (uneven sizes), affine NW 9, and an SSD request that cannot
    # be batched. Each result equals the request's mirorr_register alone, bit
    # for bit, and the budget check still covers what the slots hold.
    specs = [((64, 64, 32), (3.0, 3.0, 3.0), "rigid", 7, 3, 2),
             ((72, 56, 40), (2.5, 2.5, 3.0), "rigid", 7, 3, 2),
             ((48, 48, 48), (3.0, 3.0, 3.0), "rigid", 7, 3, 2),
             ((56, 64, 36), (2.0, 2.5, 3.0), "affine", 9, 3, 2),
             ((64, 48, 40), (2.5, 3.0, 2.5), "affine", 9, 3, 2),
             ((40, 48, 40), (3.0, 3.0, 3.0), "rigid", 7, 3, 2),
             ((60, 52, 44), (2.5, 2.5, 2.5), "affine", 9, 3, 2)]
    pairs = [phantom.make_pair(dims=d, spacing=sp, n_ellipsoids=16, seed=300 + k, model=m)
             for k, (d, sp, m, _, _, _) in enumerate(specs)]

    def reg(k, metric="normalized_correlation"):
        _, _, m, nw, ng, lev = specs[k]
        r = MirorrPyramid()
        r.SetTransformType(m)
        r.SetNhoodWidth(nw)
        r.SetNhoodGap(ng)
        r.GetPyramidScheduleTuner().SetMaxLevelNumber(lev)
        r.SetBlockMetricType(metric)
        return r

    ssd = len(pairs) - 1
    metrics = ["normalized_correlation"] * (len(pairs) - 1) + ["sum_of_squared_differences"]
    alone = []
    for k, pr in enumerate(pairs):
        r = reg(k, metrics[k])
        r.SetFixedImage(pr.fixed)
        r.SetMovingImage(pr.moving)
        alone.append(r.RunMirorr())
    with RegistrationServer(slots=2, batch=4) as srv:
        ts = [srv.submit(pr.fixed, pr.moving, reg(k, metrics[k]).params, borrow=k % 2 == 0)
              for k, pr in enumerate(pairs)]
        res = [srv.poll(t) for t in ts]
        info = srv.info()
    for k, (r, (p, c)) in enumerate(zip(res, alone)):
        assert r["status"] == capi.OK, r["error"]
        np.testing.assert_array_equal(r["params"], p, err_msg=f"request {k}")
        np.testing.assert_array_equal(r["center"], c, err_msg=f"request {k}")
    starts = {}
    for r in res:
        starts.setdefault((r["slot"], r["t_start"]), []).append(r)
    print("batches:", sorted(len(v) for v in starts.values()), "max_batch", info["max_batch"])
    assert info["batch_max"] == 4 and info["max_batch"] >= 2
    assert info["reserve_peak"] <= info["hbm_budget"]
    assert res[ssd]["status"] == capi.OK
