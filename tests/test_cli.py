"""The `mirorr` command-line drop-in, paths that need no GPU: option
handling, output-transform naming, the --fresh rule, transform inversion and
header-only (reoriented) outputs, following mirorr.cxx:282-500 and
itkMirorrPyramidWrapper.txx:596-800."""
import math
import os
import subprocess

import numpy as np
import pytest

from mirorr_amd import fileio, phantom, transforms

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "mirorr_amd", "bin", "mirorr")


@pytest.fixture(scope="module", autouse=True)
def built():
    if not os.path.exists(CLI):
        from mirorr_amd import build
        build.build()


def run(args, cwd):
    r = subprocess.run([CLI, *args], cwd=cwd, capture_output=True, text=True, timeout=120)
    return r.returncode, r.stdout, r.stderr


@pytest.fixture()
def images(tmp_path):
    p = phantom.make_pair(dims=(20, 18, 12), spacing=(2.0, 2.0, 3.0), n_ellipsoids=6, seed=2)
    fileio.write_nifti(str(tmp_path / "mov.nii.gz"), p.moving)
    fileio.write_nifti(str(tmp_path / "fix.nii.gz"), p.fixed)
    return tmp_path


def test_help_and_missing_inputs(images):
    rc, out, _ = run(["--help"], images)
    assert rc == 1 and "--nhoodwidth" in out and "--last-tfm" in out
    rc, _, err = run([], images)
    assert rc == 255 and "No moving image supplied" in err
    rc, _, err = run(["-m", "mov.nii.gz"], images)
    assert rc == 255 and "No fixed image supplied" in err
    rc, _, err = run(["none.nii", "fix.nii.gz"], images)
    assert rc == 255 and "none.nii not found" in err


@pytest.mark.parametrize("args,msg", [
    (["--reg-mode", "other"], "Unknown reg-mode"),
    (["--blockmetric", "xx"], "not a valid blockmetric"),
    (["--blockmetric", "mi"], "not supported"),
    (["-t", "quat"], "not supported"),
    (["--resampling-mode", "max-size"], "Unknown resampling type"),
    (["--resampling-mode", "other"], "Unknown resampling-mode"),
    (["--save-reoriented-moving", "a.nii", "--save-resampled-moving", "b.nii"],
     "Cannot use save-resampled-moving and save-reoriented-moving"),
    (["--bogus"], "unrecognised option"),
    (["-n"], "needs a value"),
])
def test_rejected_options(images, args, msg):
    rc, _, err = run(["mov.nii.gz", "fix.nii.gz", "-R", *args], images)
    assert rc != 0 and msg in err


def test_default_output_name_and_ignore(images):
    # No last-tfm and -R: nothing written; "ignore" never writes.
    rc, _, err = run(["mov.nii.gz", "fix.nii.gz", "-R"], images)
    assert rc == 0 and not os.path.exists(images / "fix_tfmFinal.tfm")
    rc, _, _ = run(["mov.nii.gz", "fix.nii.gz", "-R", "-l", "ignore"], images)
    assert rc == 0 and not os.path.exists(images / "ignore")


def test_do_not_register_writes_centred_initial_transform(images):
    rc, out, _ = run(["mov.nii.gz", "fix.nii.gz", "out.tfm", "-R", "-q"], images)
    assert rc == 0 and "Not running registration" in out
    params, c = fileio.read_tfm(str(images / "out.tfm"), affine=False)
    # Same grids: the centred initialisation is the identity; centre 0.
    assert np.allclose(params, 0) and np.allclose(c, 0)
    text = open(images / "out.tfm").read()
    assert text.startswith("#Insight Transform File V1.0\n#Transform 0\n"
                           "Transform: Euler3DTransform_double_3_3\n")


def test_existing_output_is_reused_unless_fresh(images):
    fileio.write_tfm(str(images / "last.tfm"), [0.1, 0, 0, 1, 2, 3], [0, 0, 0], affine=False)
    rc, out, _ = run(["mov.nii.gz", "fix.nii.gz", "last.tfm", "-R"], images)
    assert rc == 0 and "WILL be used to initialise" in out
    p, _ = fileio.read_tfm(str(images / "last.tfm"), affine=False)
    assert np.allclose(p, [0.1, 0, 0, 1, 2, 3])
    rc, out, _ = run(["mov.nii.gz", "fix.nii.gz", "last.tfm", "-R", "--fresh"], images)
    assert rc == 0 and "will NOT be used" in out
    p, _ = fileio.read_tfm(str(images / "last.tfm"), affine=False)
    assert np.allclose(p, 0)  # the centred (identity) start replaced it


def test_start_transform_inversion_round_trip(images):
    fileio.write_tfm(str(images / "s.tfm"), [0.05, -0.1, 0.2, 4, -5, 6], [1, 2, 3], affine=False)
    rc, _, _ = run(["mov.nii.gz", "fix.nii.gz", "o.tfm", "s.tfm", "-R", "-i", "-I"], images)
    assert rc == 0
    a, ca = fileio.read_tfm(str(images / "s.tfm"), affine=False)
    b, cb = fileio.read_tfm(str(images / "o.tfm"), affine=False)
    assert np.allclose(cb, 0)
    Ma, oa = transforms.matrix_offset(a, ca, False)
    Mb, ob = transforms.matrix_offset(b, cb, False)
    assert np.allclose(Ma, Mb, atol=1e-12) and np.allclose(oa, ob, atol=1e-10)


def test_unreadable_start_transform_writes_example(images):
    (images / "bad.tfm").write_text("#Insight Transform File V1.0\n#Transform 0\n"
                                    "Transform: AffineTransform_double_3_3\n"
                                    "Parameters: 2 0 0 0 1 0 0 0 1 0 0 0\nFixedParameters: 0 0 0\n")
    rc, _, err = run(["mov.nii.gz", "fix.nii.gz", "o.tfm", "bad.tfm", "-R"], images)
    assert rc == 0 and "Could not read input transform file" in err
    assert os.path.exists(images / "__tmp_example-Euler3DTransform_double_3_3.tfm")
    (images / "bsp.tfm").write_text("Transform: BSplineTransform_double_3_3\nParameters: 0\n")
    rc, _, err = run(["mov.nii.gz", "fix.nii.gz", "o.tfm", "bsp.tfm", "-R"], images)
    assert rc != 0 and "MatrixOffsetTransform" in err


def test_reoriented_moving_is_header_only(images):
    ang, t = [0.0, 0.0, math.radians(20)], [5.0, -3.0, 2.0]
    fileio.write_tfm(str(images / "r.tfm"), [*ang, *t], [0, 0, 0], affine=False)
    rc, _, _ = run(["mov.nii.gz", "fix.nii.gz", "r.tfm", "-R", "--save-reoriented-moving",
                    "rm.nii.gz", "--save-reoriented-fixed", "rf.nii"], images)
    assert rc == 0
    mov = fileio.read_nifti(str(images / "mov.nii.gz"))
    rm = fileio.read_nifti(str(images / "rm.nii.gz"))
    assert np.array_equal(rm.data, mov.data)
    M, off = transforms.matrix_offset([*ang, *t], [0, 0, 0], False)
    # Direction columns rotate with the transform, the origin moves with it.
    assert np.allclose(np.reshape(rm.direction, (3, 3)), M @ np.reshape(mov.direction, (3, 3)),
                       atol=1e-6)
    assert np.allclose(rm.origin, M @ np.asarray(mov.origin) + off, atol=1e-4)
    fix = fileio.read_nifti(str(images / "fix.nii.gz"))
    rf = fileio.read_nifti(str(images / "rf.nii"))
    Mi = M.T
    assert np.allclose(rf.origin, Mi @ (np.asarray(fix.origin) - off), atol=1e-4)


def test_reorient_shifts_cancel_in_the_written_transform(images):
    # With a start transform and no registration, the identity-space shifts
    # enter the start transform and leave the written one: the output equals
    # the start (itkMirorrPyramidWrapper.txx:626-700).
    fileio.write_tfm(str(images / "s.tfm"), [0.05, -0.1, 0.2, 4, -5, 6], [1, 2, 3], affine=False)
    for flags in (["--reorient-moving"], ["--reorient-fixed"],
                  ["--reorient-moving", "--reorient-fixed"]):
        rc, out, err = run(["mov.nii.gz", "fix.nii.gz", "o.tfm", "s.tfm", "-R", *flags], images)
        assert rc == 0, err
        a, ca = fileio.read_tfm(str(images / "s.tfm"), affine=False)
        b, cb = fileio.read_tfm(str(images / "o.tfm"), affine=False)
        Ma, oa = transforms.matrix_offset(a, ca, False)
        Mb, ob = transforms.matrix_offset(b, cb, False)
        assert np.allclose(Ma, Mb, atol=1e-12) and np.allclose(oa, ob, atol=1e-9), flags


def test_reorient_moving_with_centred_start(images):
    # The centred start on the shifted moving image, composed with the shift,
    # is the centred start on the original grids: identity for equal grids.
    rc, out, err = run(["mov.nii.gz", "fix.nii.gz", "o.tfm", "-R", "--reorient-moving"], images)
    assert rc == 0, err
    p, c = fileio.read_tfm(str(images / "o.tfm"), affine=False)
    assert np.allclose(p, 0, atol=1e-9)


def test_crop_option_parsing(images):
    for arg in ("2,3,1,1", "2x3", "(1:1)", ""):
        rc, _, err = run(["mov.nii.gz", "fix.nii.gz", "-R", "--crop", arg], images)
        assert rc == 0, (arg, err)
    rc, _, err = run(["mov.nii.gz", "fix.nii.gz", "-R", "--crop", "a,b"], images)
    assert rc != 0 and "Failed to parse" in err
    rc, _, err = run(["mov.nii.gz", "fix.nii.gz", "-R", "--crop", "10,0,0,0"], images)
    assert rc != 0 and "whole image" in err
