"""The oracle's end-to-end registration (RunMirorr level loop with classic or
symmetric StartOptimization) recovers the seeded ground truth of small
phantom pairs. These are accuracy checks of the restated algorithm, not
parity checks."""
import numpy as np
import pytest

from mirorr_amd import phantom, transforms


def _corner_error(params, center, affine, pair):
    M, off = transforms.matrix_offset(params, center, affine)
    return transforms.max_corner_distance(M, off, pair.matrix, pair.translation, pair.fixed)


@pytest.mark.parametrize("symmetric,affine", [(1, 0), (0, 0), (1, 1)])
def test_oracle_recovers_phantom_transform(oracle, symmetric, affine):
    pair = phantom.make_pair(dims=(64, 64, 32), spacing=(3.0, 3.0, 3.0), n_ellipsoids=20,
                             seed=11, model="affine" if affine else "rigid")
    rp = oracle.reg_params(symmetric=symmetric, affine=affine, nw=7, ng=3, pyr_num=2,
                           nthreads=8)
    par, c, st = oracle.register(pair.fixed, pair.moving, rp)
    assert st["levels"] == 2 and all(n > 100 for n in st["points"])
    # The fit lands within two 3-mm voxels over the field-of-view corners
    # (5 iterations per level on a coarse phantom) and removes most of the
    # initial misalignment.
    err = _corner_error(par, c, affine, pair)
    init = transforms.max_corner_distance(np.eye(3), np.zeros(3), pair.matrix,
                                          pair.translation, pair.fixed)
    assert err < 6.0 and err < init / 5.0, (err, init)


def test_oracle_variance_selection_applies_from_level_two(oracle):
    # Levels 0 and 1 use every block (itkMirorrPyramidImplement.txx:116-119):
    # with two levels the selection changes nothing; with three it thins the
    # finest level and the transform is still recovered.
    pair = phantom.make_pair(dims=(96, 96, 48), spacing=(2.0, 2.0, 2.0), n_ellipsoids=20,
                             seed=12, model="rigid")
    two = dict(symmetric=1, affine=0, nw=7, ng=3, pyr_num=2, nthreads=8)
    _, _, a = oracle.register(pair.fixed, pair.moving, oracle.reg_params(**two))
    _, _, b = oracle.register(pair.fixed, pair.moving,
                              oracle.reg_params(variance_select=1, **two))
    assert a["voxel_evals"] == b["voxel_evals"] and a["points"] == b["points"]
    three = dict(two, pyr_num=3, pyr_min_size=16)
    _, _, full = oracle.register(pair.fixed, pair.moving, oracle.reg_params(**three))
    par, c, sel = oracle.register(pair.fixed, pair.moving,
                                  oracle.reg_params(variance_select=1, **three))
    assert sel["levels"] == 3
    assert sel["voxel_evals"] < full["voxel_evals"]
    assert sel["points"][2] < full["points"][2]
    err = _corner_error(par, c, 0, pair)
    init = transforms.max_corner_distance(np.eye(3), np.zeros(3), pair.matrix,
                                          pair.translation, pair.fixed)
    assert err < 4.0 and err < init / 5.0, (err, init)


@pytest.mark.parametrize("symmetric", [1, 0])
@pytest.mark.parametrize("mode,expect", [(0, "middle"), (1, "basic"), (2, "moving"), (3, "fixed")])
def test_resampling_mode_grids(oracle, symmetric, mode, expect):
    """The level grid of each resampling mode (itkSymMirorrRegistrationMethod.txx:264-301,
    itkMirorrRegistrationMethod.txx:450-531): one level, fixed 40x36x24 and
    moving 32x44x20 voxels of 2 mm."""
    from mirorr_amd import phantom
    from mirorr_amd.phantom import Volume
    a = phantom.make_pair(dims=(40, 36, 24), spacing=(2.0, 2.0, 2.0), n_ellipsoids=8, seed=1)
    b = phantom.make_pair(dims=(32, 44, 20), spacing=(2.0, 2.0, 2.0), n_ellipsoids=8, seed=1)
    fixed, moving = a.fixed, b.moving
    rp = oracle.reg_params(symmetric=symmetric, affine=0, nw=5, ng=4, pyr_num=1, iterations=1,
                           resampling=mode)
    _, _, st = oracle.register(fixed, moving, rp)
    got = st["level_dims"][0]
    want = {"middle": (36, 40, 22), "moving": (32, 44, 20), "fixed": (40, 36, 24),
            "basic": (36, 40, 22) if symmetric else (32, 44, 20)}[expect]
    assert tuple(got) == want
