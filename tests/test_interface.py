"""The nipype interface of the reference (python/mirorr/interface.py),
restated in mirorr_amd/interface.py: command lines, checks and output lists
on the CPU; the command lines it builds are accepted by the `mirorr`
executable (a -R run needs no GPU)."""
import os
import shlex
import subprocess

import pytest

from mirorr_amd import fileio, phantom
from mirorr_amd.interface import Mirorr


@pytest.fixture()
def files(tmp_path):
    p = phantom.make_pair(dims=(20, 18, 12), spacing=(2.0, 2.0, 3.0), n_ellipsoids=6, seed=2)
    fileio.write_nifti(str(tmp_path / "mov.nii.gz"), p.moving)
    fileio.write_nifti(str(tmp_path / "fix.nii.gz"), p.fixed)
    return tmp_path


def _args(cmdline):
    return shlex.split(cmdline)[1:]


def test_minimal_cmdline_generates_transform_name(files):
    m = Mirorr(cwd=str(files), moving_image=str(files / "mov.nii.gz"),
               fixed_image=str(files / "fix.nii.gz"))
    assert _args(m.cmdline) == ["--fixed", str(files / "fix.nii.gz"),
                                "--moving", str(files / "mov.nii.gz"),
                                "--last-tfm", str(files / "rigid.tfm"),
                                "--tfm-type", "rigid"]


def test_full_cmdline_in_alphabetical_input_order(files):
    (files / "start.tfm").write_text("")
    m = Mirorr(cwd=str(files), moving_image=str(files / "mov.nii.gz"),
               fixed_image=str(files / "fix.nii.gz"), transform_type="affine",
               initial_moving_transform=str(files / "start.tfm"),
               invert_initial_moving_transform=True, invert_output_transform=False,
               transform="out.tfm", perform_registration=False, metric="CR", block_width=5,
               number_of_iterations=-4, pyramid_start=2, pyramid_end=0,
               reoriented_moving="rm.nii.gz", resampled_fixed="rf.nii.gz")
    assert _args(m.cmdline) == [
        "--blockwidth", "5", "--fixed", str(files / "fix.nii.gz"),
        "--start-tfm", str(files / "start.tfm"), "--invert-start-tfm", "--blockmetric", "cr",
        "--moving", str(files / "mov.nii.gz"), "--iterations", "-4",
        "--do-not-register", "--pyr-end", "0", "--pyr-start", "2",
        "--save-reoriented-moving", "rm.nii.gz", "--save-resampled-fixed", "rf.nii.gz",
        "--last-tfm", "out.tfm", "--tfm-type", "affine"]


def test_checks_of_the_spec(files):
    with pytest.raises(FileNotFoundError):
        Mirorr(moving_image=str(files / "none.nii"))
    with pytest.raises(ValueError, match="fixed_image"):
        Mirorr(moving_image=str(files / "mov.nii.gz")).cmdline
    with pytest.raises(ValueError, match="mutually exclusive"):
        Mirorr(moving_image=str(files / "mov.nii.gz"), fixed_image=str(files / "fix.nii.gz"),
               reoriented_fixed="a.nii", resampled_fixed="b.nii").cmdline
    with pytest.raises(ValueError):
        Mirorr(transform_type="quat")
    # "--pyr-num" has no format field: the reference's spec cannot format it.
    with pytest.raises(TypeError):
        Mirorr(moving_image=str(files / "mov.nii.gz"), fixed_image=str(files / "fix.nii.gz"),
               number_of_pyramid_levels=3).cmdline


def test_output_lists(files):
    m = Mirorr(cwd=str(files), moving_image=str(files / "mov.nii.gz"),
               fixed_image=str(files / "fix.nii.gz"), resampled_fixed="rf.nii.gz",
               resampled_moving="rm.nii.gz")
    out = m._list_outputs()
    assert out["forward_transforms"] == out["reverse_transforms"] == [str(files / "rigid.tfm")]
    assert out["forward_invert_flags"] == [False] and out["reverse_invert_flags"] == [True]
    assert out["resampled_moving"] == str(files / "rm.nii.gz")
    assert out["warped_image"] == out["resampled_moving"]
    assert out["inverse_warped_image"] == out["resampled_fixed"]
    # The reference's alias code reads the other key: one alone fails.
    m = Mirorr(cwd=str(files), moving_image=str(files / "mov.nii.gz"),
               fixed_image=str(files / "fix.nii.gz"), resampled_moving="rm.nii.gz")
    with pytest.raises(KeyError):
        m._list_outputs()


def test_cmdline_runs_without_registration(files):
    m = Mirorr(cwd=str(files), moving_image=str(files / "mov.nii.gz"),
               fixed_image=str(files / "fix.nii.gz"), perform_registration=False,
               transform="t.tfm", reoriented_moving="rm.nii.gz")
    rc, out, err, outputs = m.run(timeout=120)
    assert rc == 0, err
    assert os.path.exists(outputs["reoriented_moving"])
    assert os.path.exists(files / "t.tfm")


def test_interpolation_flag_is_rejected_like_the_reference(files):
    # The interface writes "--final-interpolater", which is neither the
    # option "--final-interpolator" nor a prefix of it: the reference's
    # option parser rejects it, and so does this executable.
    m = Mirorr(cwd=str(files), moving_image=str(files / "mov.nii.gz"),
               fixed_image=str(files / "fix.nii.gz"), perform_registration=False,
               interpolation="BSpline")
    assert "--final-interpolater bspline" in m.cmdline
    r = subprocess.run(shlex.split(m.cmdline), cwd=files, capture_output=True, text=True,
                       timeout=60)
    assert r.returncode != 0 and "unrecognised option" in r.stderr
