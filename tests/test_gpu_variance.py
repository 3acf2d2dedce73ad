"""Block-variance selection on the GPU against the oracle (through the C ABI).

The selection keeps the top `portion` of the base blocks by centred sum of
squares (ties with the cut-off kept); dropped blocks are not searched and
report the initial score -3 at their own position. The device sums in a
different order than the oracle, so a key within one float ulp of the
cut-off may land on the other side: the kept sets must agree on >= 99.9% of
the blocks, and the searched blocks meet the usual per-block bar.
"""
import numpy as np
import pytest

from mirorr_amd import transforms
from mirorr_amd.phantom import make_pair
from tests.test_gpu_bm import _compare, _matcher

pytestmark = pytest.mark.gpu


def _subset(d, m):
    return {k: d[k][m] for k in ("block_idx", "match_idx", "best", "second")}


@pytest.mark.parametrize("nw,ng,portion", [(7, 4, 0.5), (9, 3, 0.3), (6, 3, 0.5)])
def test_variance_selection_matches_oracle(oracle, nw, ng, portion):
    p = make_pair(dims=(96, 96, 48), spacing=(1.5, 1.5, 1.5), n_ellipsoids=20, seed=21)
    bm = _matcher(nw=nw, ng=ng)
    bm.SetVarianceSelection(True)
    bm.SetPortionMatchesKept(portion)
    bm.SetBaseImage(p.moving)
    bm.SetSearchImage(p.fixed)
    got = bm.get_raw()
    keep, _, nkeep = oracle.variance_select(p.moving.data, bw=4, ng=ng, portion=portion)
    gkeep = ~((got["best"] == -3.0) & np.all(got["match_idx"] == got["block_idx"], axis=1))
    agree = (gkeep == keep.astype(bool)).mean()
    print(f"blocks={keep.size} kept oracle={nkeep} gpu={gkeep.sum()} agree={agree:.6f}")
    assert agree >= 0.999
    from oracle.oracle import Geom
    g = Geom(dims=p.moving.data.shape[::-1], spacing=p.moving.spacing, origin=p.moving.origin)
    ref = oracle.bm_get_offsets(p.moving.data, None, p.fixed.data, None, g, bw=4, bg=1, nw=nw,
                                ng=ng, nthreads=8, keep=keep)
    both = gkeep & keep.astype(bool)
    assert _compare(_subset(got, both), _subset(ref, both)) >= 0.999
    dropped = ~gkeep
    assert np.all(got["best"][dropped] == -3.0) and np.all(got["second"][dropped] == -3.0)
    evals = bm.voxel_evals()
    ref_evals = oracle.candidate_count_sel(g.dims, 4, 1, nw, ng, keep) * 64
    assert abs(evals - ref_evals) <= 0.002 * ref_evals, (evals, ref_evals)


def test_variance_selection_off_and_full_portion_search_everything(oracle):
    p = make_pair(dims=(64, 64, 32), spacing=(2.0, 2.0, 2.0), n_ellipsoids=10, seed=22)
    plain = _matcher(nw=7, ng=3)
    plain.SetBaseImage(p.moving)
    plain.SetSearchImage(p.fixed)
    a = plain.get_raw()
    full = _matcher(nw=7, ng=3)
    full.SetVarianceSelection(True)
    full.SetPortionMatchesKept(1.0)
    full.SetBaseImage(p.moving)
    full.SetSearchImage(p.fixed)
    b = full.get_raw()
    for k in ("match_idx", "best", "second"):
        np.testing.assert_array_equal(a[k], b[k])
    assert plain.voxel_evals() == full.voxel_evals()


def test_registration_with_variance_selection_matches_oracle(oracle):
    from mirorr_amd.registration import MirorrPyramid
    pair = make_pair(dims=(96, 96, 48), spacing=(2.0, 2.0, 2.0), n_ellipsoids=20, seed=12,
                     model="rigid")
    rp = oracle.reg_params(symmetric=1, affine=0, nw=7, ng=3, pyr_num=3, pyr_min_size=16,
                           nthreads=8, variance_select=1)
    ref_p, ref_c, ref_st = oracle.register(pair.fixed, pair.moving, rp)
    reg = MirorrPyramid()
    reg.SetNhoodWidth(7)
    reg.SetNhoodGap(3)
    reg.SetVarianceSelection(True)
    reg.GetPyramidScheduleTuner().SetMaxLevelNumber(3)
    reg.GetPyramidScheduleTuner().SetMinLength(16)
    reg.SetFixedImage(pair.fixed)
    reg.SetMovingImage(pair.moving)
    got_p, got_c = reg.RunMirorr()
    st = reg.stats
    print("oracle", ref_st["iterations"], ref_st["points"], ref_st["voxel_evals"])
    print("gpu   ", st["iterations"], st["points"], st["voxel_evals"])
    assert st["levels"] == 3
    assert abs(st["voxel_evals"] - ref_st["voxel_evals"]) <= 0.002 * ref_st["voxel_evals"]
    M1, o1 = transforms.matrix_offset(got_p, got_c, False)
    M2, o2 = transforms.matrix_offset(ref_p, ref_c, False)
    dmm = transforms.max_corner_distance(M1, o1, M2, o2, pair.fixed)
    ddeg = transforms.rotation_angle_deg(M1, M2)
    print(f"corner distance {dmm:.3e} mm, rotation {ddeg:.3e} deg")
    assert dmm < 0.01 and ddeg < 0.01
