"""Batched registrations (mirorr_register_batch): several pairs in lockstep,
one launch of each search kernel per pass over a flat work-list of every
pair's tiles and one launch per LTS phase for all pairs. Each pair's
transform and statistics must equal mirorr_register on that pair alone, bit
for bit: pairs of different sizes and spacings (so different pyramid
schedules and level counts), masks on some, symmetric and classic methods,
rigid and affine, binary16 storage, and pairs that stop at different
iterations.
"""
import numpy as np
import pytest

from mirorr_amd import capi, phantom

pytestmark = pytest.mark.gpu


def _pairs(specs):
    out = []
    for k, (dims, sp, seed, model) in enumerate(specs):
        pair = phantom.make_pair(dims=dims, spacing=(sp, sp, sp), n_ellipsoids=16, seed=seed,
                                 model=model)
        out.append(pair)
    return out


def _reg(ctx, symmetric=True, affine=False, nw=9, levels=3, iterations=5, storage="f32",
         stop=None, vsel=False, resampling=None):
    from mirorr_amd.registration import MirorrPyramid
    reg = MirorrPyramid(ctx)
    if not symmetric:
        reg.SetRegistrationTypeClassic()
    reg.SetTransformType("affine" if affine else "rigid")
    reg.SetNhoodWidth(nw)
    reg.SetNhoodGap(3)
    reg.SetMaxIterations(iterations)
    reg.GetPyramidScheduleTuner().SetMaxLevelNumber(levels)
    reg.SetStorage(storage)
    if stop is not None:
        reg.SetStopTolerances(*stop)
    if vsel:
        reg.SetVarianceSelection(True)
    if resampling:
        reg.SetResamplingType(resampling)
    return reg


def _volumes(pairs, ctx, storage, masks):
    from mirorr_amd.registration import DeviceVolume
    F, M = [], []
    for k, pr in enumerate(pairs):
        fm = mm = None
        if masks and k % 2 == 1:
            fm = (pr.fixed.data > np.percentile(pr.fixed.data, 25)).astype(np.uint8)
            mm = (pr.moving.data > np.percentile(pr.moving.data, 25)).astype(np.uint8)
        F.append(DeviceVolume(pr.fixed, ctx, mask=fm, storage=storage))
        M.append(DeviceVolume(pr.moving, ctx, mask=mm, storage=storage))
    return F, M


def _check(pairs, *, masks=False, storage="f32", init=None, **kw):
    from mirorr_amd.blockmatcher import Context
    from mirorr_amd.registration import register_batch
    ctx_a, ctx_b = Context(0), Context(0)
    Fa, Ma = _volumes(pairs, ctx_a, storage, masks)
    Fb, Mb = _volumes(pairs, ctx_b, storage, masks)
    alone = []
    ra = _reg(ctx_a, storage=storage, **kw)
    for k, (f, m) in enumerate(zip(Fa, Ma)):
        ip, ic = (None, None) if init is None else (init[0][k], init[1][k])
        p, c = ra.RunMirorr(init_params=ip, init_center=ic, fixed_dev=f, moving_dev=m)
        alone.append((np.array(p), np.array(c), dict(ra.stats)))
    rb = _reg(ctx_b, storage=storage, **kw)
    for rep in range(2):  # the second run reuses the batch's buffers
        got = register_batch(rb, Fb, Mb, *(init or (None, None)))
        for k, ((p, c), st) in enumerate(zip(got, rb.batch_stats)):
            pa, ca, sa = alone[k]
            np.testing.assert_array_equal(p, pa, err_msg=f"pair {k} run {rep}")
            np.testing.assert_array_equal(c, ca, err_msg=f"pair {k} run {rep}")
            for key in ("levels", "iterations", "points", "inner_iterations", "voxel_evals",
                        "rescored_blocks", "level_dims"):
                assert np.array_equal(np.asarray(st[key]), np.asarray(sa[key])), (k, key)
    print("levels", [a[2]["levels"] for a in alone], "iterations",
          [a[2]["iterations"] for a in alone])
    for v in Fa + Ma + Fb + Mb:
        v.close()
    ctx_a.close()
    ctx_b.close()
    return alone


UNEVEN = [((64, 64, 32), 2.0, 1, "rigid"), ((96, 80, 48), 1.5, 2, "rigid"),
          ((40, 48, 40), 3.0, 3, "rigid"), ((72, 64, 56), 2.0, 4, "rigid")]
UNEVEN_AFF = [((64, 64, 32), 2.0, 5, "affine"), ((96, 80, 48), 1.5, 6, "affine"),
              ((56, 48, 40), 2.5, 7, "affine")]


def test_uneven_symmetric_rigid_batch_equals_each_pair_alone():
    alone = _check(_pairs(UNEVEN), masks=True)
    # the schedules differ, so pairs join the lockstep at different levels
    assert len({a[2]["levels"] for a in alone}) > 1


def test_uneven_symmetric_affine_batch_equals_each_pair_alone():
    _check(_pairs(UNEVEN_AFF), affine=True)


def test_classic_batch_equals_each_pair_alone():
    _check(_pairs(UNEVEN[:3]), symmetric=False, masks=True)
    _check(_pairs(UNEVEN_AFF[:2]), symmetric=False, affine=True)


def test_binary16_batch_equals_each_pair_alone():
    _check(_pairs(UNEVEN_AFF), affine=True, storage="f16", nw=7)


def test_stop_rules_and_halved_iterations_in_a_batch():
    # the reference's stop rules at their defaults (pairs stop at different
    # iterations) and negative iterations (halved per level)
    _check(_pairs(UNEVEN), iterations=-8, masks=True)


def test_start_transforms_variance_selection_and_resampling_modes_in_a_batch():
    pairs = _pairs(UNEVEN[:3])
    rng = np.random.default_rng(5)
    init_p = [np.concatenate([rng.uniform(-0.03, 0.03, 3), rng.uniform(-2, 2, 3)])
              for _ in pairs]
    init_c = [rng.uniform(-5, 5, 3) for _ in pairs]
    _check(pairs, init=(init_p, init_c))
    # variance selection acts from a pair's third level on
    big = _pairs([((128, 128, 64), 1.5, 9, "rigid"), ((144, 112, 96), 1.2, 10, "rigid")])
    alone = _check(big, vsel=True, masks=True)
    assert any(a[2]["levels"] >= 3 for a in alone)
    _check(pairs, symmetric=False, resampling="fixed")
    _check(pairs, resampling="basic", affine=False)


def test_single_pair_and_empty_batches():
    from mirorr_amd.blockmatcher import Context
    from mirorr_amd.registration import register_batch
    _check(_pairs(UNEVEN[1:2]))
    ctx = Context(0)
    reg = _reg(ctx)
    assert register_batch(reg, [], []) == []
    ctx.close()


def test_batch_rejects_what_it_cannot_batch():
    from mirorr_amd.blockmatcher import Context
    from mirorr_amd.registration import DeviceVolume, register_batch
    pr = _pairs(UNEVEN[:1])[0]
    ctx = Context(0)
    f = DeviceVolume(pr.fixed, ctx)
    m = DeviceVolume(pr.moving, ctx)
    m16 = DeviceVolume(pr.moving, ctx, storage="f16")
    reg = _reg(ctx)
    reg.SetBlockMetricType("sum_of_squared_differences")
    with pytest.raises(capi.MirorrError, match="NCC"):
        register_batch(reg, [f], [m])
    reg = _reg(ctx)
    with pytest.raises(capi.MirorrError, match="storage"):
        register_batch(reg, [f], [m16])
    reg = _reg(ctx)
    reg.SetNhoodWidth(15)
    with pytest.raises(capi.MirorrError, match="tiled"):
        register_batch(reg, [f], [m])
    reg = _reg(ctx)
    ctx.set_profiling(True)
    with pytest.raises(capi.MirorrError, match="time"):
        register_batch(reg, [f], [m])
    ctx.set_profiling(False)
    for v in (f, m, m16):
        v.close()
    ctx.close()


def test_batch_mixing_large_and_small_lts_runs():
    # The 256x256x128 pair's full-resolution LTS run has more than 262,144
    # pairs (eight per moment thread), the other pair's runs fewer (four per
    # thread): one iteration launches both moment kernels, each for its own
    # pairs, and each pair still equals its own registration.
    _check(_pairs([((256, 256, 128), 1.0, 21, "affine"), ((96, 80, 48), 1.5, 22, "affine")]),
           affine=True, iterations=2)
