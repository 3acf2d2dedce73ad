"""Pins the CPU oracle's block matcher (itkBlockMatcher.txx) against closed
forms, invariants and a pure-Python brute force on tiny volumes."""
import numpy as np
import pytest

from tests import bm_bruteforce as bf


def _rand(shape, seed, lo=0.0, hi=100.0):
    return np.random.default_rng(seed).uniform(lo, hi, shape).astype(np.float32)


def test_ncc_linear_relation_is_one(oracle):
    x = _rand((6, 6, 6), 0)
    for a, b in [(2.5, 10.0), (-0.7, 300.0), (1.0, 0.0)]:
        y = (a * x.astype(np.float64) + b).astype(np.float32)
        r2 = oracle.ncc_pair(x, None, y, None, (1, 1, 1), (1, 1, 1), 4)
        assert abs(r2 - 1.0) < 1e-6


def test_ncc_constant_block_is_minus_two(oracle):
    x = _rand((6, 6, 6), 1)
    y = np.full((6, 6, 6), 7.0, np.float32)
    assert oracle.ncc_pair(x, None, y, None, (0, 0, 0), (2, 2, 2), 4) == -2.0
    assert oracle.ncc_pair(y, None, x, None, (0, 0, 0), (2, 2, 2), 4) == -2.0


def test_ncc_fully_masked_is_minus_two(oracle):
    x = _rand((6, 6, 6), 2)
    m = np.zeros((6, 6, 6), np.uint8)
    assert oracle.ncc_pair(x, m, x, None, (0, 0, 0), (0, 0, 0), 4) == -2.0


def test_ncc_affine_intensity_invariance(oracle):
    x = _rand((8, 8, 8), 3)
    y = _rand((8, 8, 8), 4)
    r0 = oracle.ncc_pair(x, None, y, None, (1, 2, 3), (3, 2, 1), 4)
    y2 = (-3.0 * y.astype(np.float64) + 55.0).astype(np.float32)
    r1 = oracle.ncc_pair(x, None, y2, None, (1, 2, 3), (3, 2, 1), 4)
    assert abs(r0 - r1) < 1e-6
    # swapping the roles of the two blocks leaves r^2 unchanged
    r2 = oracle.ncc_pair(y, None, x, None, (3, 2, 1), (1, 2, 3), 4)
    assert abs(r0 - r2) < 1e-12


@pytest.mark.parametrize("seed", [5, 6])
def test_ncc_masked_matches_bruteforce(oracle, seed):
    rng = np.random.default_rng(seed)
    x = _rand((7, 7, 7), seed)
    y = _rand((7, 7, 7), seed + 10)
    mx = (rng.uniform(size=x.shape) > 0.3).astype(np.uint8) * 255
    my = (rng.uniform(size=x.shape) > 0.3).astype(np.uint8) * 200
    for b, s in [((0, 0, 0), (3, 3, 3)), ((2, 1, 0), (0, 3, 2))]:
        ref = bf.ncc(x, mx, y, my, b, s, 4)
        got = oracle.ncc_pair(x, mx, y, my, b, s, 4)
        assert got == ref


def test_block_grid_closed_form(oracle):
    assert oracle.block_grid((128, 128, 64), 4, 4) == ((32, 32, 16), (0, 0, 0))
    # 256x256x128 with NG=3: n = (L-4)//3+1, centred start (L - (3(n-1)+4))//2
    assert oracle.block_grid((256, 256, 128), 4, 3) == ((85, 85, 42), (0, 0, 0))
    assert oracle.block_grid((13, 10, 9), 4, 2) == ((5, 4, 3), (0, 0, 0))
    assert oracle.block_grid((130, 12, 8), 4, 4) == ((32, 3, 2), (1, 0, 0))
    assert oracle.block_grid((14, 10, 9), 4, 3) == ((4, 3, 2), (0, 0, 1))


def test_candidate_count_config1(oracle):
    # 128x128x64, BW=4, NG=4, NW=7, BG=1: per axis sum of clamped ranges.
    def axis(L, nb):
        tot = 0
        for i in range(nb):
            p = 4 * i
            tot += min(p + 3, L - 4) - max(p - 3, 0) + 1
        return tot
    expect = axis(128, 32) * axis(128, 32) * axis(64, 16)
    assert oracle.candidate_count((128, 128, 64), 4, 1, 7, 4) == expect == 5037544


PARAMS = [
    dict(bw=4, bg=1, nw=3, ng=2),
    dict(bw=4, bg=1, nw=5, ng=3),
    dict(bw=3, bg=2, nw=4, ng=3),   # even NW, odd BW, BG>1 clipping at edges
    dict(bw=2, bg=3, nw=3, ng=2),
]


@pytest.mark.parametrize("pi", range(len(PARAMS)))
@pytest.mark.parametrize("masked", [False, True])
def test_get_offsets_matches_bruteforce(oracle, pi, masked):
    p = PARAMS[pi]
    shape = (9, 10, 12)  # (nz, ny, nx)
    rng = np.random.default_rng(100 + pi)
    base = _rand(shape, 200 + pi)
    # search: a smoothed copy shifted by one voxel plus noise (structured matches)
    search = np.roll(base, shift=(1, -1, 1), axis=(0, 1, 2)) * 0.5 + _rand(shape, 300 + pi, 0, 5)
    search = search.astype(np.float32)
    bm = sm = None
    if masked:
        bm = (rng.uniform(size=shape) > 0.15).astype(np.uint8) * 255
        sm = (rng.uniform(size=shape) > 0.15).astype(np.uint8) * 255
    from oracle.oracle import Geom
    g = Geom(dims=shape[::-1])
    got = oracle.bm_get_offsets(base, bm, search, sm, g, nthreads=0, **p)
    ref = bf.get_offsets(base, bm, search, sm, p["bw"], p["bg"], p["nw"], p["ng"])
    assert len(ref) == got["best"].shape[0]
    for i, (bidx, midx, best, second) in enumerate(ref):
        assert tuple(got["block_idx"][i]) == bidx
        assert tuple(got["match_idx"][i]) == midx
        assert got["best"][i] == best
        assert got["second"][i] == second


def test_threaded_equals_single(oracle):
    from oracle.oracle import Geom
    shape = (20, 24, 28)
    base = _rand(shape, 7)
    search = np.roll(base, 2, axis=2) + _rand(shape, 8, 0, 20)
    search = search.astype(np.float32)
    g = Geom(dims=shape[::-1], spacing=(1.5, 1.2, 2.0), origin=(-3.0, 4.0, 10.0))
    one = oracle.bm_get_offsets(base, None, search, None, g, bw=4, bg=1, nw=5, ng=3)
    for T in (2, 3, 7):
        many = oracle.bm_get_offsets(base, None, search, None, g, bw=4, bg=1, nw=5, ng=3,
                                     nthreads=T)
        for k in ("block_idx", "match_idx", "best", "second"):
            np.testing.assert_array_equal(one[k], many[k])
        # kept lists are a permutation (thread-slab concatenation order)
        a = np.lexsort(one["base_mm"].T)
        b = np.lexsort(many["base_mm"].T)
        np.testing.assert_array_equal(one["base_mm"][a], many["base_mm"][b])
        np.testing.assert_array_equal(one["match_mm"][a], many["match_mm"][b])


def test_exact_tie_first_in_scan_order_wins_and_is_culled(oracle):
    from oracle.oracle import Geom
    shape = (12, 12, 12)
    base = _rand(shape, 9)
    search = _rand(shape, 10, 0, 1)
    blk = base[4:8, 4:8, 4:8]
    # two identical, non-overlapping copies of the block at x-offsets -2 and +2
    search[4:8, 4:8, 2:6] = blk
    search[4:8, 4:8, 6:10] = blk
    g = Geom(dims=shape[::-1])
    out = oracle.bm_get_offsets(base, None, search, None, g, bw=4, bg=1, nw=5, ng=4)
    nb = out["nblocks"]
    # block at start (4,4,4) is the centre one of the 3x3x3 grid at stride 4
    i = 1 + nb[0] * (1 + nb[1] * 1)
    assert tuple(out["block_idx"][i]) == (4, 4, 4)
    assert tuple(out["match_idx"][i]) == (2, 4, 4)  # x=2 comes first in scan order
    assert out["best"][i] == out["second"][i]
    # culled: its point is absent from the kept list (centre = idx + BW/2)
    kept = {tuple(r) for r in out["base_mm"]}
    assert (6.0, 6.0, 6.0) not in kept


def test_integer_shift_recovered_with_mm_points(oracle):
    from oracle.oracle import Geom
    shape = (24, 24, 24)
    base = _rand(shape, 11)
    search = np.roll(base, shift=(0, 1, -2), axis=(0, 1, 2)).astype(np.float32)
    sp = (1.5, 2.0, 1.0)
    org = (-10.0, 5.0, 2.5)
    g = Geom(dims=shape[::-1], spacing=sp, origin=org)
    out = oracle.bm_get_offsets(base, None, search, None, g, bw=4, bg=1, nw=7, ng=4)
    bi, mi = out["block_idx"], out["match_idx"]
    interior = np.all((bi >= 3) & (bi <= 24 - 4 - 3), axis=1)
    d = mi[interior] - bi[interior]
    assert np.all(d == np.array([-2, 1, 0]))
    assert np.allclose(out["best"][interior], 1.0, atol=1e-6)
    # mm conversion: origin + spacing*idx + spacing*BW/2 (identity direction)
    for pb, pm in zip(out["base_mm"], out["match_mm"]):
        ib = (pb - np.array(org)) / np.array(sp) - 2.0
        assert np.allclose(ib, np.round(ib))


# ---- block-variance selection ----------------------------------------------

def _checker_blocks(amps, bw=4):
    """Volume whose blocks (ng = bw, grid from the corner) are +/-a checkerboards,
    so each block's centred sum of squares is exactly bw^3 * a^2."""
    nb = amps.shape  # (z, y, x) blocks
    vol = np.zeros((nb[0] * bw, nb[1] * bw, nb[2] * bw), np.float32)
    z, y, x = np.indices((bw, bw, bw))
    sign = np.where((x + y + z) % 2 == 0, 1.0, -1.0).astype(np.float32)
    for bz in range(nb[0]):
        for by in range(nb[1]):
            for bx in range(nb[2]):
                vol[bz * bw:(bz + 1) * bw, by * bw:(by + 1) * bw, bx * bw:(bx + 1) * bw] = \
                    100.0 + amps[bz, by, bx] * sign
    return vol


@pytest.mark.parametrize("portion", [0.1, 0.25, 0.5, 0.9])
def test_variance_select_keeps_top_portion(oracle, portion):
    rng = np.random.default_rng(11)
    amps = rng.permutation(np.arange(1, 4 * 5 * 6 + 1)).reshape(4, 5, 6).astype(np.float32)
    vol = _checker_blocks(amps)
    keep, keys, n = oracle.variance_select(vol, bw=4, ng=4, portion=portion)
    flat = amps.reshape(-1)  # z, y, x raster = block raster order (x fastest)
    np.testing.assert_allclose(keys, 64.0 * flat ** 2, rtol=1e-6)
    K = int(np.ceil(portion * flat.size))
    expect = np.zeros(flat.size, np.uint8)
    expect[np.argsort(-flat)[:K]] = 1
    np.testing.assert_array_equal(keep, expect)
    assert n == K


def test_variance_select_ties_and_limits(oracle):
    amps = np.ones((2, 2, 3), np.float32)
    amps[0, 0, 0] = 5.0
    vol = _checker_blocks(amps)
    keep, _, n = oracle.variance_select(vol, bw=4, ng=4, portion=0.5)
    assert n == amps.size and keep.all()  # cut-off ties with 11 blocks: all kept
    keep, _, n = oracle.variance_select(vol, bw=4, ng=4, portion=1.0 / 24)
    assert n == 1 and keep[0] == 1  # at least one block, the largest
    keep, _, n = oracle.variance_select(vol, bw=4, ng=4, portion=1.0)
    assert keep.all()


def test_variance_key_is_population_variance(oracle):
    rng = np.random.default_rng(5)
    vol = rng.normal(300.0, 50.0, (12, 16, 20)).astype(np.float32)
    keep, keys, _ = oracle.variance_select(vol, bw=4, ng=3, portion=0.5)
    nb, st = oracle.block_grid((20, 16, 12), 4, 3)
    i = 0
    for z in range(nb[2]):
        for y in range(nb[1]):
            for x in range(nb[0]):
                blk = vol[st[2] + 3 * z:st[2] + 3 * z + 4, st[1] + 3 * y:st[1] + 3 * y + 4,
                          st[0] + 3 * x:st[0] + 3 * x + 4].astype(np.float64)
                assert keys[i] == pytest.approx(64.0 * blk.var(), rel=1e-6)
                i += 1


def test_selected_search_skips_dropped_blocks(oracle):
    rng = np.random.default_rng(9)
    base = rng.normal(0.0, 1.0, (16, 20, 24)).astype(np.float32)
    search = np.roll(base, 1, axis=2).copy()
    g = oracle.Geom(dims=(24, 20, 16))
    keep, _, n = oracle.variance_select(base, bw=4, ng=3, portion=0.5)
    full = oracle.bm_get_offsets(base, None, search, None, g, bw=4, bg=1, nw=5, ng=3)
    sel = oracle.bm_get_offsets(base, None, search, None, g, bw=4, bg=1, nw=5, ng=3, keep=keep)
    on = keep.astype(bool)
    np.testing.assert_array_equal(sel["match_idx"][on], full["match_idx"][on])
    np.testing.assert_array_equal(sel["best"][on], full["best"][on])
    assert np.all(sel["best"][~on] == -3.0)
    np.testing.assert_array_equal(sel["match_idx"][~on], sel["block_idx"][~on])
    assert sel["base_mm"].shape[0] <= n
    cc = oracle.candidate_count_sel(g.dims, 4, 1, 5, 3, keep)
    assert cc < oracle.candidate_count(g.dims, 4, 1, 5, 3)
