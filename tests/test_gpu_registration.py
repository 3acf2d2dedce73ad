"""Parity of the device resampler, pyramid, LTS and full registration with
the CPU oracle (through the C ABI)."""
import numpy as np
import pytest

from mirorr_amd import phantom, transforms
from mirorr_amd.capi import Geom as CGeom
from oracle.oracle import Geom

pytestmark = pytest.mark.gpu


def _geoms(dims, spacing=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0)):
    return (CGeom.make(dims, spacing, origin), Geom(dims=dims, spacing=spacing, origin=origin))


def test_resample_matches_oracle(oracle):
    from mirorr_amd.registration import resample
    rng = np.random.default_rng(0)
    a = rng.uniform(0, 1000, (20, 24, 28)).astype(np.float32)
    gi_c, gi_o = _geoms((28, 24, 20), (1.5, 1.2, 2.0), (-20.0, -14.0, -19.0))
    go_c, go_o = _geoms((30, 26, 18), (1.4, 1.3, 2.1), (-21.0, -16.0, -18.0))
    M = phantom.euler_zxy(0.1, -0.05, 0.15) @ np.diag([1.05, 0.97, 1.02])
    off = np.array([1.3, -2.2, 0.7])
    got, gm = resample(a, gi_c, M, off, go_c, want_mask=True)
    ref, rm = oracle.resample(a, None, gi_o, M.ravel(), off, go_o, want_mask=True)
    np.testing.assert_allclose(got, ref, rtol=1e-6, atol=1e-3)
    assert np.mean(gm != rm) < 1e-3  # boundary points may round either way


@pytest.mark.parametrize("sched", [[[2, 2, 2], [1, 1, 1]], [[4, 4, 2], [2, 2, 2], [1, 1, 1]],
                                   [[2, 2, 1], [1, 1, 1]]])
def test_pyramid_matches_oracle(oracle, sched):
    from mirorr_amd.registration import pyramid_level
    rng = np.random.default_rng(1)
    a = rng.uniform(0, 100, (16, 32, 40)).astype(np.float32)
    gc, go = _geoms((40, 32, 16), (1.0, 1.5, 2.0), (3.0, -4.0, 5.0))
    rows = len(sched)
    # oracle recursion
    prev, pg = a, go
    levels = {}
    for lvl in range(rows - 1, -1, -1):
        lg = oracle.pyramid_level_geom(go, sched[lvl])
        ratio = sched[lvl] if lvl == rows - 1 else [sched[lvl][k] // sched[lvl + 1][k]
                                                    for k in range(3)]
        prev = oracle.pyramid_step(prev, pg, ratio, lg)
        pg = lg
        levels[lvl] = (prev, lg)
    for lvl in range(rows):
        got, gg = pyramid_level(a, gc, sched, lvl)
        ref, rg = levels[lvl]
        assert tuple(gg.dims) == tuple(rg.dims)
        np.testing.assert_allclose(tuple(gg.origin), rg.origin, atol=1e-12)
        np.testing.assert_allclose(got, ref, rtol=1e-6, atol=1e-4)


def _pairs(seed, n=3000, affine=False, outliers=0.3):
    rng = np.random.default_rng(seed)
    X = rng.uniform(-80, 80, (n, 3))
    if affine:
        A = phantom.euler_zxy(0.08, -0.05, 0.1) @ np.diag([1.04, 0.97, 1.02])
    else:
        A = phantom.euler_zxy(0.08, -0.05, 0.1)
    Y = X @ A.T + np.array([3.0, -2.0, 5.0]) + rng.normal(0, 0.5, X.shape)
    bad = rng.choice(n, int(outliers * n), replace=False)
    Y[bad] += rng.uniform(-30, 30, (len(bad), 3))
    return X, Y


@pytest.mark.parametrize("affine", [0, 1])
def test_lts_classic_matches_oracle(oracle, affine):
    from mirorr_amd.registration import lts_fit
    X, Y = _pairs(2, affine=bool(affine))
    n = 12 if affine else 6
    p0 = np.zeros(n)
    if affine:
        p0[[0, 4, 8]] = 1.0
    c0 = np.array([1.0, -2.0, 0.5])
    rp, rc, rit = oracle.lts_classic(affine, X, Y, p0, c0)
    gp, gc, _, _, git = lts_fit(False, bool(affine), X, Y, p0, c0)
    assert git == rit
    np.testing.assert_allclose(gp, rp, atol=1e-9)


@pytest.mark.parametrize("affine", [0, 1])
def test_lts_symmetric_matches_oracle(oracle, affine):
    from mirorr_amd.registration import lts_fit
    X, Y = _pairs(3, affine=bool(affine))
    rf, rb, rp, rit = oracle.lts_symmetric(affine, X, Y)
    gp, _, gf, gb, git = lts_fit(True, bool(affine), X, Y)
    assert git == rit
    np.testing.assert_allclose(gf, rf, atol=1e-9)
    np.testing.assert_allclose(gb, rb, atol=1e-9)
    np.testing.assert_allclose(gp, rp, atol=1e-9)


def test_lts_symmetric_exact_on_clean_affine():
    from mirorr_amd.registration import lts_fit
    X, Y = _pairs(4, affine=True, outliers=0.0)
    rng = np.random.default_rng(4)
    A = phantom.euler_zxy(0.08, -0.05, 0.1) @ np.diag([1.04, 0.97, 1.02])
    Y = X @ A.T + np.array([3.0, -2.0, 5.0])
    _, _, f, b, _ = lts_fit(True, True, X, Y)
    T = np.vstack([f, [0, 0, 0, 1]]) @ np.linalg.inv(np.vstack([b, [0, 0, 0, 1]]))
    np.testing.assert_allclose(T[:3, :3], A, atol=1e-9)


@pytest.mark.parametrize("symmetric,affine", [(1, 0), (1, 1), (0, 0), (0, 1)])
def test_registration_matches_oracle(oracle, symmetric, affine):
    from mirorr_amd.registration import MirorrPyramid
    pair = phantom.make_pair(dims=(64, 64, 32), spacing=(3.0, 3.0, 3.0), n_ellipsoids=20,
                             seed=11, model="affine" if affine else "rigid")
    rp = oracle.reg_params(symmetric=symmetric, affine=affine, nw=7, ng=3, pyr_num=2,
                           nthreads=8)
    ref_p, ref_c, ref_st = oracle.register(pair.fixed, pair.moving, rp)
    reg = MirorrPyramid()
    if not symmetric:
        reg.SetRegistrationTypeClassic()
    reg.SetTransformType("affine" if affine else "rigid")
    reg.SetNhoodWidth(7)
    reg.SetNhoodGap(3)
    reg.GetPyramidScheduleTuner().SetMaxLevelNumber(2)
    reg.SetFixedImage(pair.fixed)
    reg.SetMovingImage(pair.moving)
    got_p, got_c = reg.RunMirorr()
    st = reg.stats
    print("oracle", ref_st["iterations"], ref_st["points"], ref_st["inner_iterations"])
    print("gpu   ", st["iterations"], st["points"], st["inner_iterations"], st["seconds"])
    assert st["voxel_evals"] > 0
    M1, o1 = transforms.matrix_offset(got_p, got_c, affine)
    M2, o2 = transforms.matrix_offset(ref_p, ref_c, affine)
    dmm = transforms.max_corner_distance(M1, o1, M2, o2, pair.fixed)
    ddeg = transforms.rotation_angle_deg(M1, M2)
    print(f"corner distance {dmm:.3e} mm, rotation {ddeg:.3e} deg")
    assert dmm < 0.01 and ddeg < 0.01



@pytest.mark.parametrize("symmetric,mode", [(1, "middle"), (1, "moving"), (1, "fixed"),
                                            (0, "middle"), (0, "basic"), (0, "moving"),
                                            (0, "fixed")])
def test_resampling_modes_match_oracle(oracle, symmetric, mode):
    from mirorr_amd import capi
    from mirorr_amd.registration import MirorrPyramid
    # One phantom pair; the moving image cropped to another field of view so
    # that the grids of the modes differ.
    pair = phantom.make_pair(dims=(64, 64, 32), spacing=(3.0, 3.0, 3.0), n_ellipsoids=20,
                             seed=21)
    fixed = pair.fixed
    m = pair.moving
    sub = np.ascontiguousarray(m.data[2:30, 0:64, 4:60])
    org = (m.origin[0] + 4 * 3.0, m.origin[1], m.origin[2] + 2 * 3.0)
    moving = phantom.Volume(sub, m.spacing, org)
    # One iteration per level: over five, float-level differences flip an
    # LTS stop decision or a block at the cull threshold in some modes of this
    # pair and the runs drift apart by tenths of a millimetre (the
    # sensitivity documented for the workload pairs); the grids and
    # resampling a mode changes are all exercised by the first iteration.
    rp = oracle.reg_params(symmetric=symmetric, affine=0, nw=7, ng=3, pyr_num=2, nthreads=8,
                           resampling=capi.RESAMPLING_NAMES[mode], iterations=1)
    ref_p, ref_c, ref_st = oracle.register(fixed, moving, rp)
    reg = MirorrPyramid()
    if not symmetric:
        reg.SetRegistrationTypeClassic()
    reg.SetTransformType("rigid")
    reg.SetNhoodWidth(7)
    reg.SetNhoodGap(3)
    reg.SetResamplingType(mode)
    reg.SetMaxIterations(1)
    reg.GetPyramidScheduleTuner().SetMaxLevelNumber(2)
    reg.SetFixedImage(fixed)
    reg.SetMovingImage(moving)
    print("resampling", symmetric, mode, flush=True)
    got_p, got_c = reg.RunMirorr()
    assert reg.stats["level_dims"] == ref_st["level_dims"]
    M1, o1 = transforms.matrix_offset(got_p, got_c, False)
    M2, o2 = transforms.matrix_offset(ref_p, ref_c, False)
    dmm = transforms.max_corner_distance(M1, o1, M2, o2, fixed)
    print(f"{mode}: {ref_st['level_dims']} corner distance {dmm:.3e} mm; points oracle "
          f"{ref_st['points']} gpu {reg.stats['points']}, iterations {ref_st['iterations']} "
          f"{reg.stats['iterations']}")
    assert dmm < 0.01


@pytest.mark.parametrize("affine", [0, 1])
def test_overlapped_passes_match_serial_passes(affine):
    """The symmetric iteration's two block-matching passes run on two streams
    of the context, except while search kernels are timed, when they run one
    after the other. Both orders give bit-identical registrations: the passes
    write disjoint buffers and every reduction runs in a fixed order."""
    from mirorr_amd.blockmatcher import Context
    from mirorr_amd.registration import MirorrPyramid
    pair = phantom.make_pair(dims=(96, 80, 64), spacing=(2.0, 2.0, 2.5), n_ellipsoids=20,
                             seed=21 + affine, model="affine" if affine else "rigid")
    ctx = Context()
    out = {}
    for name, timed in (("overlapped", False), ("serial", True)):
        ctx.set_profiling(ctx.PROF_FAST if timed else False)
        reg = MirorrPyramid(ctx)
        reg.SetTransformType("affine" if affine else "rigid")
        reg.SetNhoodWidth(7)
        reg.GetPyramidScheduleTuner().SetMaxLevelNumber(3)
        reg.SetFixedImage(pair.fixed)
        reg.SetMovingImage(pair.moving)
        print("registration,", name, flush=True)
        out[name] = reg.RunMirorr() + (reg.stats["iterations"],)
    ctx.set_profiling(False)
    n_timed, _, _ = ctx.profile("ncc_tile_fast")
    assert n_timed > 0  # the serial run did time its search launches
    np.testing.assert_array_equal(out["overlapped"][0], out["serial"][0])
    np.testing.assert_array_equal(out["overlapped"][1], out["serial"][1])
    assert out["overlapped"][2] == out["serial"][2]


@pytest.mark.parametrize("symmetric,affine", [(1, 1), (0, 0)])
def test_long_lts_runs_match_the_oracle_and_fixed_chunks(oracle, symmetric, affine):
    """A tight LTS threshold (1e-9) makes the inner runs long (up to the
    reference's 100 iterations), so the host queues each run's first chunk
    sized by the previous run's inner iterations and then chunks of 2 and 4.
    The registration must match the oracle's (same outer iterations and
    points, transform within 1e-5 mm) and equal, bit for bit, the same
    registration with iteration graphs on, where every first chunk is 4."""
    from mirorr_amd.blockmatcher import Context
    from mirorr_amd.registration import DeviceVolume, MirorrPyramid
    pair = phantom.make_pair(dims=(64, 64, 32), spacing=(3.0, 3.0, 3.0), n_ellipsoids=20,
                             seed=12, model="affine" if affine else "rigid")
    rp = oracle.reg_params(symmetric=symmetric, affine=affine, nw=7, ng=3, pyr_num=2,
                           nthreads=8, lts_tolerance=1e-9)
    ref_p, ref_c, ref_st = oracle.register(pair.fixed, pair.moving, rp)
    runs = []
    for graphs in (False, True):
        ctx = Context(0)
        ctx.iteration_graphs = graphs
        reg = MirorrPyramid(ctx)
        if not symmetric:
            reg.SetRegistrationTypeClassic()
        reg.SetTransformType("affine" if affine else "rigid")
        reg.SetNhoodWidth(7)
        reg.SetNhoodGap(3)
        reg.GetPyramidScheduleTuner().SetMaxLevelNumber(2)
        reg.SetStopTolerances(1e-3, 1e-9)
        f = DeviceVolume(pair.fixed, ctx)
        m = DeviceVolume(pair.moving, ctx)
        p, c = reg.RunMirorr(fixed_dev=f, moving_dev=m)
        runs.append((np.array(p), np.array(c), dict(reg.stats)))
        f.close()
        m.close()
        ctx.close()
    (pa, ca, sa), (pb, cb, sb) = runs
    print("oracle", ref_st["iterations"], ref_st["points"], ref_st["inner_iterations"])
    print("gpu   ", sa["iterations"], sa["points"], sa["inner_iterations"])
    # long runs: several chunks per run on average
    assert sa["inner_iterations"] > 6 * sum(sa["iterations"][:sa["levels"]])
    np.testing.assert_array_equal(pa, pb)
    np.testing.assert_array_equal(ca, cb)
    assert sa["inner_iterations"] == sb["inner_iterations"]
    levels = ref_st["levels"]
    assert list(sa["iterations"][:levels]) == list(ref_st["iterations"][:levels])
    assert list(sa["points"][:levels]) == list(ref_st["points"][:levels])
    M1, o1 = transforms.matrix_offset(pa, ca, affine)
    M2, o2 = transforms.matrix_offset(ref_p, ref_c, affine)
    dmm = transforms.max_corner_distance(M1, o1, M2, o2, pair.fixed)
    print(f"corner distance {dmm:.3e} mm")
    assert dmm < 1e-5
