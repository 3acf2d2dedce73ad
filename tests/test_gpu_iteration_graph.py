"""Iteration graphs (mirorr_ctx_set_iteration_graphs): an outer iteration of
a level up to the LTS's first chunk (resampling, both block-matching passes,
the LTS start, its first four inner iterations and the finish) is captured
as a HIP graph at the level's second iteration, kept on the context and
replayed while the host values its launches read are unchanged. The
replayed launches are the eager ones with the same arguments, so every
registration must give bit-identical transforms and statistics with the
graphs on and off: symmetric and classic methods, rigid and affine, masks,
variance selection (one stream), binary16 storage, and the outer stop rule
switched off so that every level replays its graph.
"""
import time

import numpy as np
import pytest

from mirorr_amd import phantom

pytestmark = pytest.mark.gpu


def _run(pair, graphs, *, symmetric=True, affine=False, nw=9, levels=2, iterations=5,
         stop=None, masks=False, vsel=False, storage="f32", ctx=None, repeat=1):
    from mirorr_amd.registration import DeviceVolume, MirorrPyramid
    reg = MirorrPyramid(ctx)
    reg.ctx.iteration_graphs = graphs
    assert reg.ctx.iteration_graphs == graphs
    if not symmetric:
        reg.SetRegistrationTypeClassic()
    reg.SetTransformType("affine" if affine else "rigid")
    reg.SetNhoodWidth(nw)
    reg.SetNhoodGap(3)
    reg.SetMaxIterations(iterations)
    reg.GetPyramidScheduleTuner().SetMaxLevelNumber(levels)
    reg.SetStorage(storage)
    if vsel:
        reg.SetVarianceSelection(True)
    if stop is not None:
        reg.SetStopTolerances(*stop)
    fm = mm = None
    if masks:
        fm = (pair.fixed.data > np.percentile(pair.fixed.data, 30)).astype(np.uint8)
        mm = (pair.moving.data > np.percentile(pair.moving.data, 30)).astype(np.uint8)
    f = DeviceVolume(pair.fixed, reg.ctx, mask=fm, storage=storage)
    m = DeviceVolume(pair.moving, reg.ctx, mask=mm, storage=storage)
    out = []
    for _ in range(repeat):
        t0 = time.perf_counter()
        p, c = reg.RunMirorr(fixed_dev=f, moving_dev=m)
        out.append((np.array(p), np.array(c), dict(reg.stats), time.perf_counter() - t0))
    f.close()
    m.close()
    return out


def _same(a, b):
    pa, ca, sa, _ = a
    pb, cb, sb, _ = b
    np.testing.assert_array_equal(pa, pb)
    np.testing.assert_array_equal(ca, cb)
    for k in ("iterations", "points", "inner_iterations", "voxel_evals", "rescored_blocks",
              "level_dims"):
        assert np.array_equal(np.asarray(sa[k]), np.asarray(sb[k])), (k, sa[k], sb[k])


CASES = {
    "symmetric_affine": dict(symmetric=True, affine=True),
    "symmetric_rigid_masks_nw13": dict(symmetric=True, masks=True, nw=13),
    "classic_rigid": dict(symmetric=False),
    "classic_affine_masks": dict(symmetric=False, affine=True, masks=True),
    "variance_selection": dict(symmetric=True, vsel=True, levels=3),
    "binary16": dict(symmetric=True, affine=True, storage="f16"),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_graph_replay_is_bit_identical(name):
    kw = CASES[name]
    pair = phantom.make_pair(dims=(96, 96, 48), spacing=(2.0, 2.0, 2.0), n_ellipsoids=20, seed=11,
                             model="affine" if kw.get("affine") else "rigid")
    print(f"case {name}", flush=True)
    # stop rule off: every level runs all five iterations, three of them replays
    eager = _run(pair, False, stop=(0.0, 1e-3), **kw)[0]
    graph = _run(pair, True, stop=(0.0, 1e-3), **kw)[0]
    assert all(n == 5 for n in eager[2]["iterations"][:eager[2]["levels"]])
    _same(eager, graph)


def test_graph_replay_under_the_reference_stop_rule_and_repeated_runs():
    # The graphs are cached per level on the context and replayed from the
    # first iteration while their key matches: repeated runs, a second pair
    # of the same size (new volumes, the same level buffers) and a pinned
    # tile shape (a new key: captured again) all stay bit-identical.
    from mirorr_amd import capi
    from mirorr_amd.blockmatcher import Context
    pair = phantom.make_pair(dims=(128, 128, 64), spacing=(1.5, 1.5, 1.5), n_ellipsoids=30,
                             seed=3, model="affine")
    other = phantom.make_pair(dims=(128, 128, 64), spacing=(1.5, 1.5, 1.5), n_ellipsoids=30,
                              seed=8, model="affine")
    ce, cg = Context(0), Context(0)
    eager = _run(pair, False, affine=True, levels=3, ctx=ce, repeat=3)
    graph = _run(pair, True, affine=True, levels=3, ctx=cg, repeat=3)
    for g in graph:
        _same(eager[0], g)
    for e in eager[1:]:
        _same(eager[0], e)
    _same(_run(other, False, affine=True, levels=3, ctx=ce)[0],
          _run(other, True, affine=True, levels=3, ctx=cg)[0])
    try:
        assert capi.lib().mirorr_tile_force(9, 1) == capi.OK
        _same(eager[0], _run(pair, True, affine=True, levels=3, ctx=cg, repeat=2)[1])
    finally:
        capi.lib().mirorr_tile_force(9, -1)
    te = min(x[3] for x in eager[1:])
    tg = min(x[3] for x in graph[1:])
    print(f"128x128x64 affine, 3 levels: eager {1e3 * te:.2f} ms, graphs {1e3 * tg:.2f} ms; "
          f"iterations {eager[0][2]['iterations'][:3]}")
    ce.close()
    cg.close()
