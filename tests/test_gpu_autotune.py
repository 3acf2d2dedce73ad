"""Tile shapes of the NCC search and their autotuner (mirorr_tile_force,
mirorr_tile_autotune_report). Every shape scores every candidate with the
same operations in the same order, so every shape gives the same records,
bit for bit, on unmasked and on masked images; the autotuner times the
shapes that fit on the first search of a grid and reports its choice."""
import numpy as np
import pytest

from mirorr_amd import capi, phantom
from mirorr_amd.phantom import Volume

pytestmark = pytest.mark.gpu


def _search(base, search, bmask, smask, nw, ng=3):
    from mirorr_amd.blockmatcher import BlockMatcher
    bm = BlockMatcher()
    bm.SetNhoodWidth(nw)
    bm.SetNhoodGap(ng)
    bm.SetBaseImage(base)
    bm.SetSearchImage(search)
    if bmask is not None:
        bm.SetBaseMask(bmask)
        bm.SetSearchMask(smask)
    return bm.get_raw()


@pytest.mark.parametrize("nw", [5, 7, 9, 11, 13])
def test_every_tile_shape_gives_the_same_records(oracle, nw):
    pair = phantom.make_pair(dims=(80, 72, 48), spacing=(2.0, 2.0, 2.0), n_ellipsoids=16,
                             seed=21, model="affine")
    g = oracle.geom_of(pair.moving)
    R = phantom.euler_zxy(0.06, -0.04, 0.05)
    t = np.array([6.1, -4.3, 3.2])
    xb, mb = oracle.resample(pair.moving.data, None, g, R.ravel(), t, g, want_mask=True)
    xs, ms = oracle.resample(pair.fixed.data, None, g, R.T.ravel(), -t, g, want_mask=True)
    cases = [(pair.moving, pair.fixed, None, None),
             (Volume(xb, pair.moving.spacing, pair.moving.origin),
              Volume(xs, pair.moving.spacing, pair.moving.origin), mb, ms)]
    L = capi.lib()
    try:
        for base, search, bmask, smask in cases:
            ref = None
            for k in range(3):
                assert L.mirorr_tile_force(nw, k) == capi.OK
                got = _search(base, search, bmask, smask, nw)
                if ref is None:
                    ref = got
                    continue
                for key in ("best", "second", "match_idx"):
                    np.testing.assert_array_equal(got[key], ref[key])
        assert L.mirorr_tile_force(nw, 3) == capi.ERR_ARG
    finally:
        L.mirorr_tile_force(nw, -1)


def test_autotuner_times_the_shapes_and_reports_its_choice():
    # a grid no other test uses, so the first search tunes it
    pair = phantom.make_pair(dims=(92, 84, 60), spacing=(2.0, 2.0, 2.0), n_ellipsoids=12, seed=22)
    _search(pair.moving, pair.fixed, None, None, 9)
    report = capi.tile_autotune_report()
    print(report)
    line = [l for l in report.splitlines() if "nw 9 ng 3 f32 grid 92x84x60" in l]
    assert len(line) == 1
    assert line[0].count(" us") == 3  # three shapes timed
    again = _search(pair.moving, pair.fixed, None, None, 9)
    assert capi.tile_autotune_report().count("grid 92x84x60") == 1  # chosen once
    assert again["best"].shape[0] > 0
