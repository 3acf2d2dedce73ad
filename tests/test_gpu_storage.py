"""Binary16 storage of volumes, pyramid levels and half-way images
(mirorr_volume_create_storage, mirorr_bm_params.storage): the search widens
the encoded values to float where it stages them, so

* 12-bit integer images (the CT-like and MR-like pair) are stored exactly:
  the f16 search equals the f32 search bit for bit;
* on float images the only change is the rounding of the stored values to
  binary16: per block against the FP64 oracle on the float images, and the
  transforms of whole registrations against the oracle's. A registration's
  path through its capped iterations amplifies any perturbation of its
  inputs, so the transform bar is a quarter of the finest voxel (0.25 mm at
  1 mm, 1 mm at 4 mm voxels) and 0.05 deg, with the distance to the seeded
  truth reported beside it.
"""
import time

import numpy as np
import pytest

from mirorr_amd import capi, phantom, transforms
from mirorr_amd.phantom import Volume

pytestmark = pytest.mark.gpu


def _matcher(storage, nw=9, ng=3):
    from mirorr_amd.blockmatcher import BlockMatcher
    bm = BlockMatcher()
    bm.SetNhoodWidth(nw)
    bm.SetNhoodGap(ng)
    bm.SetStorage(storage)
    return bm


def _raw(storage, base, search, nw=9, ng=3):
    bm = _matcher(storage, nw, ng)
    bm.SetBaseImage(base)
    bm.SetSearchImage(search)
    return bm.get_raw()


def test_12bit_integer_images_are_stored_exactly():
    pair = phantom.make_ct_mr_pair(dims=(128, 128, 64), spacing=(4.0, 4.0, 4.0), n_ellipsoids=30,
                                   seed=4)
    assert np.all(pair.fixed.data == np.rint(pair.fixed.data))
    a = _raw("f32", pair.moving, pair.fixed, nw=13)
    b = _raw("f16", pair.moving, pair.fixed, nw=13)
    for k in ("best", "second", "match_idx"):
        np.testing.assert_array_equal(a[k], b[k])


def test_block_search_on_float_images_against_the_fp64_oracle(oracle):
    # Float phantom intensities: binary16 keeps 11 significant bits of the
    # encoded values (about half an intensity unit at the phantom's scale),
    # which moves the NCC^2 of low-contrast blocks most. Bar: the best
    # candidate of >= 99% of the blocks whose oracle best-second gap exceeds
    # 1e-3 is the oracle's; NCC^2 within 1e-2 of the oracle's double sums on
    # the float images at every block and within 5e-3 at 99% of them.
    pair = phantom.make_pair(dims=(96, 96, 48), spacing=(2.0, 2.0, 2.0), n_ellipsoids=20, seed=5,
                             model="affine")
    g = oracle.geom_of(pair.moving)
    ref = oracle.bm_get_offsets(pair.moving.data, None, pair.fixed.data, None, g, bw=4, bg=1,
                                nw=9, ng=3, nthreads=8)
    got = _raw("f16", pair.moving, pair.fixed)
    same = np.all(got["match_idx"] == ref["match_idx"], axis=1)
    clear = np.abs(ref["best"] - ref["second"]) > 1e-3
    db = np.abs(got["best"].astype(np.float64) - ref["best"])
    print(f"f16 search: identical best {same.mean():.4f} (clear blocks {same[clear].mean():.4f}), "
          f"max |dNCC^2| {db.max():.2e}, p99 {np.percentile(db, 99):.2e}")
    assert same[clear].mean() >= 0.99
    assert db.max() <= 1e-2 and np.percentile(db, 99) <= 5e-3


def _register(pair, storage, nw, ng, levels, affine=False, iterations=0, ctx=None):
    from mirorr_amd.registration import DeviceVolume, MirorrPyramid
    reg = MirorrPyramid(ctx)
    reg.SetTransformType("affine" if affine else "rigid")
    reg.SetNhoodWidth(nw)
    reg.SetNhoodGap(ng)
    if iterations:
        reg.SetMaxIterations(iterations)
    reg.GetPyramidScheduleTuner().SetMaxLevelNumber(levels)
    reg.SetStorage(storage)
    t = time.perf_counter()
    f = DeviceVolume(pair.fixed, reg.ctx, storage=storage)
    m = DeviceVolume(pair.moving, reg.ctx, storage=storage)
    upload = time.perf_counter() - t
    p, c = reg.RunMirorr(fixed_dev=f, moving_dev=m)
    return p, c, reg.stats, upload, f.info()


def test_ctmr_reduced_registration_within_a_quarter_voxel_of_the_oracle(oracle):
    # 128x128x64 at 4 mm, 2 levels run to the reference's stop rule.
    pair = phantom.make_ct_mr_pair(dims=(128, 128, 64), spacing=(4.0, 4.0, 4.0), n_ellipsoids=30,
                                   seed=4)
    rp = oracle.reg_params(symmetric=1, affine=0, nw=13, ng=3, pyr_num=2, iterations=20,
                           nthreads=8)
    ref_p, ref_c, ref_st = oracle.register(pair.fixed, pair.moving, rp)
    p, c, st, _, info = _register(pair, "f16", 13, 3, 2, iterations=20)
    assert info["storage"] == capi.STORAGE_F16
    M1, o1 = transforms.matrix_offset(p, c, False)
    M2, o2 = transforms.matrix_offset(ref_p, ref_c, False)
    dmm = transforms.max_corner_distance(M1, o1, M2, o2, pair.fixed)
    ddeg = transforms.rotation_angle_deg(M1, M2)
    e1 = transforms.max_corner_distance(M1, o1, pair.matrix, pair.translation, pair.fixed)
    e2 = transforms.max_corner_distance(M2, o2, pair.matrix, pair.translation, pair.fixed)
    p32, c32, st32, _, _ = _register(pair, "f32", 13, 3, 2, iterations=20)
    M3, o3 = transforms.matrix_offset(p32, c32, False)
    e3 = transforms.max_corner_distance(M3, o3, pair.matrix, pair.translation, pair.fixed)
    print(f"f16 vs oracle: {dmm:.4f} mm {ddeg:.4f} deg; iterations {st['iterations']} "
          f"(oracle {ref_st['iterations']}); vs truth: f16 {e1:.3f} mm, oracle {e2:.3f} mm, "
          f"f32 {e3:.3f} mm; points f16 {st['points']} f32 {st32['points']}")
    assert dmm < 1.0 and ddeg < 0.05


def test_pair256_affine_registration_within_a_quarter_voxel_of_the_oracle(oracle):
    # Float phantom intensities: the levels and half-way images round to
    # binary16.
    pair = phantom.make_pair(dims=(256, 256, 128), spacing=(1.0, 1.0, 1.0), n_ellipsoids=60,
                             seed=2, model="affine")
    rp = oracle.reg_params(symmetric=1, affine=1, nw=9, ng=3, pyr_num=3, iterations=5,
                           nthreads=16)
    ref_p, ref_c, ref_st = oracle.register(pair.fixed, pair.moving, rp)
    p, c, st, _, _ = _register(pair, "f16", 9, 3, 3, affine=True, iterations=5)
    M1, o1 = transforms.matrix_offset(p, c, True)
    M2, o2 = transforms.matrix_offset(ref_p, ref_c, True)
    dmm = transforms.max_corner_distance(M1, o1, M2, o2, pair.fixed)
    ddeg = transforms.rotation_angle_deg(M1, M2)
    e1 = transforms.max_corner_distance(M1, o1, pair.matrix, pair.translation, pair.fixed)
    e2 = transforms.max_corner_distance(M2, o2, pair.matrix, pair.translation, pair.fixed)
    print(f"f16 vs oracle: {dmm:.4f} mm {ddeg:.4f} deg; iterations {st['iterations']} "
          f"(oracle {ref_st['iterations']}); vs truth: f16 {e1:.3f} mm, oracle {e2:.3f} mm; "
          f"points f16 {st['points']} oracle {ref_st['points']}")
    assert dmm < 0.25 and ddeg < 0.05


def test_ctmr_full_size_recovers_with_half_the_volume_memory():
    from mirorr_amd.registration import MirorrPyramid
    from mirorr_amd.server import RegistrationServer
    pair = phantom.make_ct_mr_pair(n_ellipsoids=24, seed=3)
    r = MirorrPyramid()
    r.SetTransformType("rigid")
    r.SetNhoodWidth(13)
    r.SetNhoodGap(3)
    r.GetPyramidScheduleTuner().SetMaxLevelNumber(4)
    res = {}
    with RegistrationServer(slots=1) as srv:
        for storage in ("f32", "f16"):
            res[storage] = srv.poll(srv.submit(pair.fixed, pair.moving, r.params,
                                               storage=storage))
    init = transforms.max_corner_distance(np.eye(3), np.zeros(3), pair.matrix, pair.translation,
                                          pair.fixed)
    for storage, x in res.items():
        assert x["status"] == capi.OK, x["error"]
        M, off = transforms.matrix_offset(x["params"], x["center"], False)
        err = transforms.max_corner_distance(M, off, pair.matrix, pair.translation, pair.fixed)
        print(f"{storage}: error {err:.2f} mm (initial {init:.1f}), bound "
              f"{x['bytes_bound'] / 2**20:.0f} MiB, {1e3 * (x['t_end'] - x['t_start']):.0f} ms")
        assert err < 3.0 and err < init / 5.0
    # the volumes, pyramid levels and half-way images are half the bytes
    # (window statistics, block data and the pyramid's double scratch stay)
    assert res["f16"]["bytes_bound"] < 0.9 * res["f32"]["bytes_bound"]


def test_upload_bytes_halve():
    from mirorr_amd.registration import DeviceVolume
    pair = phantom.make_pair(dims=(256, 256, 128), spacing=(1.0, 1.0, 1.0), n_ellipsoids=8,
                             seed=1)
    t = {}
    for storage in ("f32", "f16", "f32", "f16"):
        t0 = time.perf_counter()
        v = DeviceVolume(pair.fixed, storage=storage)
        t[storage] = time.perf_counter() - t0
        v.close()
    print(f"upload of a 256x256x128 volume: f32 {1e3 * t['f32']:.1f} ms, "
          f"f16 {1e3 * t['f16']:.1f} ms (host encoding included)")


def test_mixed_storage_and_unsupported_searches_raise():
    from mirorr_amd.registration import DeviceVolume, MirorrPyramid
    pair = phantom.make_pair(dims=(48, 48, 32), spacing=(3.0, 3.0, 3.0), n_ellipsoids=8, seed=2)
    reg = MirorrPyramid()
    reg.GetPyramidScheduleTuner().SetMaxLevelNumber(1)
    f = DeviceVolume(pair.fixed, reg.ctx, storage="f16")
    m = DeviceVolume(pair.moving, reg.ctx, storage="f32")
    with pytest.raises(capi.MirorrError, match="different storage"):
        reg.RunMirorr(fixed_dev=f, moving_dev=m)
    m16 = DeviceVolume(pair.moving, reg.ctx, storage="f16")
    reg.SetBlockMetricType("sum_of_squared_differences")
    with pytest.raises(capi.MirorrError, match="half-precision"):
        reg.RunMirorr(fixed_dev=f, moving_dev=m16)
