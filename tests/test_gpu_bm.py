"""Parity of the HIP NCC block search with the CPU oracle (through the C ABI).

Bar (per block, block raster order): identical best candidate for >= 99.9%
of blocks, and |best - best_oracle|, |second - second_oracle| <= 1e-5 (the
north-star tolerance for NCC^2 values, fp32 device sums against the oracle's
double sums). Blocks whose best differs must be near-ties in the oracle
(|best - second| <= 1e-4: the reference culls them anyway).
"""
import numpy as np
import pytest

from mirorr_amd.phantom import Volume, make_pair

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _matcher(bw=4, bg=1, nw=7, ng=3):
    from mirorr_amd.blockmatcher import BlockMatcher
    bm = BlockMatcher()
    bm.SetBlockWidth(bw)
    bm.SetBlockGap(bg)
    bm.SetNhoodWidth(nw)
    bm.SetNhoodGap(ng)
    return bm


def _compare(gpu, ref, min_frac=0.999):
    """Per-block parity. Blocks whose oracle best and second are within 1e-4
    are ties the reference culls (itkBlockMatcher.txx:894); their argmax is
    decided by rounding, so the identical-best bar applies to the others."""
    n = ref["best"].shape[0]
    assert gpu["best"].shape[0] == n
    np.testing.assert_array_equal(gpu["block_idx"], ref["block_idx"])
    same = np.all(gpu["match_idx"] == ref["match_idx"], axis=1)
    gap = np.abs(ref["best"].astype(np.float64) - ref["second"])
    clear = gap > 1e-4 + 2 * TOL
    frac = same[clear].mean() if clear.any() else 1.0
    print(f"blocks={n} identical-best={same.mean():.6f} clear={clear.sum()} "
          f"identical-among-clear={frac:.6f}")
    if not np.all(same | ~clear):
        bad = np.where(~same & clear)[0][:10]
        for i in bad:
            print("block", i, gpu["block_idx"][i], "gpu", gpu["match_idx"][i], gpu["best"][i],
                  gpu["second"][i], "oracle", ref["match_idx"][i], ref["best"][i],
                  ref["second"][i])
    assert frac >= min_frac, f"only {frac:.6f} of clear blocks share the best candidate"
    db = np.abs(gpu["best"].astype(np.float64) - ref["best"])
    ds = np.abs(gpu["second"].astype(np.float64) - ref["second"])
    assert db.max(initial=0) <= TOL, db.max()
    assert ds.max(initial=0) <= TOL, ds.max()
    return same.mean()


def _oracle_raw(oracle, vb, mb, vs, ms, bw, bg, nw, ng):
    from oracle.oracle import Geom
    g = Geom(dims=vb.data.shape[::-1], spacing=vb.spacing, origin=vb.origin)
    return oracle.bm_get_offsets(vb.data, mb, vs.data, ms, g, bw=bw, bg=bg, nw=nw, ng=ng,
                                 nthreads=8)


def _vol(a, spacing=(1.0, 1.0, 1.0)):
    a = np.ascontiguousarray(a, np.float32)
    org = tuple(-0.5 * (d - 1) * s for d, s in zip(a.shape[::-1], spacing))
    return Volume(a, spacing, org)


def _smooth_random(shape, seed):
    rng = np.random.default_rng(seed)
    a = rng.normal(0, 1, shape)
    for ax in range(3):
        a = (np.roll(a, 1, ax) + 2 * a + np.roll(a, -1, ax)) / 4
    return (500 + 200 * a).astype(np.float32)


@pytest.mark.parametrize("nw,ng", [(5, 2), (7, 3), (7, 4), (9, 3), (11, 3), (13, 4)])
def test_tiled_unmasked_matches_oracle(oracle, nw, ng):
    shape = (26, 30, 34)
    base = _smooth_random(shape, nw)
    search = (np.roll(base, (1, -2, 1), (0, 1, 2)) * -0.7 + 900 +
              np.random.default_rng(nw + 1).normal(0, 5, shape)).astype(np.float32)
    vb, vs = _vol(base), _vol(search)
    bm = _matcher(nw=nw, ng=ng)
    bm.SetBaseImage(vb)
    bm.SetSearchImage(vs)
    _compare(bm.get_raw(), _oracle_raw(oracle, vb, None, vs, None, 4, 1, nw, ng))


def _fov_mask(shape, normal, offset):
    z, y, x = np.meshgrid(*[np.arange(s) for s in shape], indexing="ij")
    v = normal[0] * x + normal[1] * y + normal[2] * z + offset
    return (v > 0).astype(np.uint8) * 255


@pytest.mark.parametrize("nw,ng", [(5, 4), (7, 4), (9, 3)])
def test_masked_matches_oracle(oracle, nw, ng):
    shape = (28, 32, 36)
    base = _smooth_random(shape, 40 + nw)
    search = (np.roll(base, (0, 1, -1), (0, 1, 2)) * 1.3 - 50 +
              np.random.default_rng(nw).normal(0, 5, shape)).astype(np.float32)
    mb = _fov_mask(shape, (1.0, 0.3, 0.1), -6.0)
    ms = _fov_mask(shape, (-0.2, 1.0, 0.5), -5.0)
    vb, vs = _vol(base), _vol(search)
    bm = _matcher(nw=nw, ng=ng)
    bm.SetBaseImage(vb)
    bm.SetSearchImage(vs)
    bm.SetBaseMask(mb)
    bm.SetSearchMask(ms)
    _compare(bm.get_raw(), _oracle_raw(oracle, vb, mb, vs, ms, 4, 1, nw, ng))


@pytest.mark.parametrize("nw,ng", [(5, 2), (9, 3), (13, 4)])
def test_all_off_base_blocks_match_oracle(oracle, nw, ng):
    # Blocks whose base mask is off everywhere score -2 at every candidate
    # (itkBlockMatcher.txx:96-97); the device writes their record without a
    # search. A slab of off planes plus a partial search mask covers them
    # next to partly masked and fully on blocks.
    shape = (30, 32, 36)
    base = _smooth_random(shape, 70 + nw)
    search = (np.roll(base, (1, 0, 1), (0, 1, 2)) * 0.8 + 20 +
              np.random.default_rng(nw).normal(0, 4, shape)).astype(np.float32)
    mb = np.full(shape, 255, np.uint8)
    mb[:11] = 0
    mb[:, :, 30:] = 0
    ms = _fov_mask(shape, (0.2, -1.0, 0.3), 20.0)
    vb, vs = _vol(base), _vol(search)
    bm = _matcher(nw=nw, ng=ng)
    bm.SetBaseImage(vb)
    bm.SetSearchImage(vs)
    bm.SetBaseMask(mb)
    bm.SetSearchMask(ms)
    ref = _oracle_raw(oracle, vb, mb, vs, ms, 4, 1, nw, ng)
    got = bm.get_raw()
    off = np.array([not mb[z:z + 4, y:y + 4, x:x + 4].any() for x, y, z in ref["block_idx"]])
    assert off.sum() > 20 and (~off).sum() > 20
    np.testing.assert_array_equal(got["match_idx"][off], ref["match_idx"][off])
    np.testing.assert_array_equal(got["best"][off], ref["best"][off])
    np.testing.assert_array_equal(got["second"][off], ref["second"][off])
    assert np.all(ref["best"][off] == -2.0)
    _compare(got, ref)


def test_all_off_base_single_candidate(oracle):
    # One block, one candidate: best -2, second stays at the initial -3.
    shape = (4, 4, 4)
    base = _smooth_random(shape, 5)
    vb, vs = _vol(base), _vol(base * 2.0)
    mb = np.zeros(shape, np.uint8)
    bm = _matcher(nw=5, ng=2)
    bm.SetBaseImage(vb)
    bm.SetSearchImage(vs)
    bm.SetBaseMask(mb)
    got = bm.get_raw()
    ref = _oracle_raw(oracle, vb, mb, vs, None, 4, 1, 5, 2)
    assert ref["best"].tolist() == [-2.0] and ref["second"].tolist() == [-3.0]
    np.testing.assert_array_equal(got["match_idx"], ref["match_idx"])
    np.testing.assert_array_equal(got["best"], ref["best"])
    np.testing.assert_array_equal(got["second"], ref["second"])


@pytest.mark.parametrize("bw,bg,nw,ng", [(3, 2, 4, 3), (4, 1, 3, 2), (5, 1, 5, 4), (4, 2, 5, 3)])
def test_generic_kernel_matches_oracle(oracle, bw, bg, nw, ng):
    shape = (18, 20, 22)
    base = _smooth_random(shape, 7 * bw + nw)
    search = (np.roll(base, (1, 0, -1), (0, 1, 2)) +
              np.random.default_rng(3).normal(0, 20, shape)).astype(np.float32)
    mb = _fov_mask(shape, (1.0, 0.0, 0.2), -3.0)
    vb, vs = _vol(base), _vol(search)
    bm = _matcher(bw, bg, nw, ng)
    bm.SetBaseImage(vb)
    bm.SetSearchImage(vs)
    bm.SetBaseMask(mb)
    ref = _oracle_raw(oracle, vb, mb, vs, None, bw, bg, nw, ng)
    got = bm.get_raw()
    # double accumulation on both sides: identical candidates and float scores
    np.testing.assert_array_equal(got["match_idx"], ref["match_idx"])
    np.testing.assert_allclose(got["best"], ref["best"], atol=1e-6)
    np.testing.assert_allclose(got["second"], ref["second"], atol=1e-6)


def test_phantom_level_parity_and_offsets(oracle):
    p = make_pair(dims=(96, 96, 48), spacing=(1.5, 1.5, 1.5), n_ellipsoids=20, seed=3)
    bm = _matcher(nw=7, ng=4)
    bm.SetBaseImage(p.moving)
    bm.SetSearchImage(p.fixed)
    ref = _oracle_raw(oracle, p.moving, None, p.fixed, None, 4, 1, 7, 4)
    assert _compare(bm.get_raw(), ref) >= 0.999
    bmm, mmm = bm.GetOffsets()
    # Blocks next to the filter's thresholds are re-scored in double, so the
    # kept set and the matched points are the oracle's, in block raster order
    # (the oracle's threaded list is a permutation of it).
    key = lambda a: np.lexsort(np.round(a, 6).T[::-1])
    rb, rm = ref["base_mm"][key(ref["base_mm"])], ref["match_mm"][key(ref["base_mm"])]
    gb, gm = bmm[key(bmm)], mmm[key(bmm)]
    assert gb.shape == rb.shape
    np.testing.assert_allclose(gb, rb, atol=1e-9)
    np.testing.assert_allclose(gm, rm, atol=1e-9)


def _periodic(shape, period, seed):
    cell = _smooth_random(shape[:2] + (period,), seed)
    return np.ascontiguousarray(np.tile(cell, (1, 1, shape[2] // period)))


@pytest.mark.parametrize("masked", [False, True])
def test_searched_exact_ties_match_oracle(oracle, masked):
    # Base = search, periodic along x with period 4: with NW 9 a block's
    # candidates 4 voxels apart see identical windows, so interior blocks
    # have exact ties at their best score. The reference's float score buffer
    # (itkBlockMatcher.txx:823-833) keeps the first of the tied candidates in
    # scan order unless the double score rounds down to float, in which case
    # each later tie replaces it; either way second = best and the filter
    # culls the block (|best - second| <= 1e-4). A partial base mask sends
    # the blocks through the masked kernel.
    shape = (20, 24, 32)
    img = _periodic(shape, 4, 31)
    vb, vs = _vol(img), _vol(img)
    mb = None
    bm = _matcher(nw=9, ng=3)
    bm.SetBaseImage(vb)
    bm.SetSearchImage(vs)
    if masked:
        mb = np.full(shape, 255, np.uint8)
        mb[:, :, 1::7] = 0  # one x column in seven: most blocks are partly masked
        bm.SetBaseMask(mb)
    got = bm.get_raw()
    ref = _oracle_raw(oracle, vb, mb, vs, None, 4, 1, 9, 3)
    tie = (ref["best"] == ref["second"]) & (ref["best"] > 0.5)
    first = ref["match_idx"][:, 0] < ref["block_idx"][:, 0]
    print(f"tied blocks {tie.sum()} of {tie.size}; first-in-scan-order winners "
          f"{(tie & first).sum()}")
    assert tie.sum() > 50
    np.testing.assert_array_equal(got["match_idx"][tie], ref["match_idx"][tie])
    np.testing.assert_array_equal(got["best"][tie], ref["best"][tie])
    np.testing.assert_array_equal(got["second"][tie], ref["second"][tie])
    # every tied block is culled
    assert not _kept_any(got["best"][tie], got["second"][tie])
    _compare(got, ref)


@pytest.mark.parametrize("masked", [False, True])
def test_rescored_flat_neighbourhoods_match_oracle(oracle, masked):
    # Base = search = a linear ramp (integer values, exact in float): every
    # candidate window is the block plus a constant, so all 729 candidates of
    # an interior block score NCC^2 = 1 exactly, in float and in double, on
    # any subset of voxels. The exact float ties are re-scored: all 729
    # intervals reach the cut, so the double phase scores every candidate.
    # The records must equal the oracle's (the same sums in the same order,
    # the first tied candidate in scan order) and the tied blocks are culled. A partial base mask sends the blocks through the
    # masked kernel and the masked sums.
    shape = (24, 28, 32)
    z, y, x = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shape], indexing="ij")
    img = 3.0 * x + 2.0 * y + z + 100.0
    vb, vs = _vol(img), _vol(img)
    mb = None
    bm = _matcher(nw=9, ng=3)
    bm.SetBaseImage(vb)
    bm.SetSearchImage(vs)
    if masked:
        mb = np.full(shape, 255, np.uint8)
        mb[:, :, 1::7] = 0
        bm.SetBaseMask(mb)
    got = bm.get_raw()
    ref = _oracle_raw(oracle, vb, mb, vs, None, 4, 1, 9, 3)
    tie = (ref["best"] == ref["second"]) & (ref["best"] > 0.5)
    print(f"blocks {tie.size}, tied at NCC^2 = 1: {tie.sum()}")
    assert tie.sum() > 0.9 * tie.size
    np.testing.assert_array_equal(got["match_idx"][tie], ref["match_idx"][tie])
    np.testing.assert_array_equal(got["best"][tie], ref["best"][tie])
    np.testing.assert_array_equal(got["second"][tie], ref["second"][tie])
    assert not _kept_any(got["best"][tie], got["second"][tie])
    _compare(got, ref)


def _kept_any(best, second):
    b = np.asarray(best, np.float64)
    s = np.asarray(second, np.float64)
    return bool(np.any((b > 1e-7) & (np.abs(b - s) > 1e-4)))


def test_swap_equals_fresh_reverse_matcher(oracle):
    p = make_pair(dims=(64, 64, 32), spacing=(2.0, 2.0, 2.0), n_ellipsoids=10, seed=5)
    bm = _matcher(nw=5, ng=3)
    bm.SetBaseImage(p.moving)
    bm.SetSearchImage(p.fixed)
    bm.get_raw()
    bm.SwapImageOrder()
    swapped = bm.get_raw()
    fresh = _matcher(nw=5, ng=3)
    fresh.SetBaseImage(p.fixed)
    fresh.SetSearchImage(p.moving)
    f = fresh.get_raw()
    for k in ("match_idx", "best", "second"):
        np.testing.assert_array_equal(swapped[k], f[k])


def test_no_match_raises_algo_error():
    from mirorr_amd.capi import AlgoError
    a = np.full((16, 16, 16), 3.0, np.float32)
    bm = _matcher(nw=3, ng=4)
    bm.SetBaseImage(_vol(a))
    bm.SetSearchImage(_vol(a))
    raw = bm.get_raw()
    assert np.all(raw["best"] == -2.0)
    with pytest.raises(AlgoError):
        bm.GetOffsets()


def test_fully_masked_raises_algo_error():
    from mirorr_amd.capi import AlgoError
    base = _smooth_random((16, 16, 16), 1)
    bm = _matcher(nw=5, ng=4)
    bm.SetBaseImage(_vol(base))
    bm.SetSearchImage(_vol(base))
    bm.SetBaseMask(np.zeros(base.shape, np.uint8))
    with pytest.raises(AlgoError):
        bm.GetOffsets()


def test_image_smaller_than_block_is_an_error():
    from mirorr_amd.capi import AlgoError
    a = np.ones((2, 8, 8), np.float32)
    bm = _matcher(nw=3, ng=2)
    bm.SetBaseImage(_vol(a))
    bm.SetSearchImage(_vol(a))
    with pytest.raises(AlgoError):
        bm.get_raw()
