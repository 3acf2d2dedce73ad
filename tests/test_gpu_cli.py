"""End to end through the `mirorr` command line on the GPU: NIfTI in,
transform file and resampled image out, compared with the CPU oracle run on
the same (ARI-reoriented) volumes and with the oracle's final resampling."""
import os
import subprocess

import numpy as np
import pytest

from mirorr_amd import fileio, phantom, transforms
from oracle import final_resample as fr

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "mirorr_amd", "bin", "mirorr")


class _G:
    def __init__(self, v):
        self.dims, self.spacing, self.origin, self.direction = (v.dims, v.spacing, v.origin,
                                                                v.direction)


@pytest.mark.parametrize("affine", [0, 1])
def test_cli_matches_oracle(tmp_path, oracle, affine):
    pair = phantom.make_pair(dims=(64, 64, 32), spacing=(3.0, 3.0, 3.0), n_ellipsoids=20,
                             seed=11, model="affine" if affine else "rigid")
    fileio.write_nifti(str(tmp_path / "mov.nii.gz"), pair.moving)
    fileio.write_nifti(str(tmp_path / "fix.nii"), pair.fixed)
    args = [CLI, "-m", "mov.nii.gz", "-f", "fix.nii", "-l", "out.tfm", "-t",
            "affine" if affine else "rigid", "-d", "2", "-c", "0", "--fresh",
            "--save-moving", "res.nii.gz"]
    print("cli", " ".join(args[1:]), flush=True)
    r = subprocess.run(args, cwd=tmp_path, capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stderr + r.stdout[-2000:]
    params, c = fileio.read_tfm(str(tmp_path / "out.tfm"), affine=bool(affine))
    assert np.allclose(c, 0)
    # The oracle on the volumes as the command line sees them: read back
    # from the files and reoriented to ARI.
    mov = fileio.reorient_ari(fileio.read_nifti(str(tmp_path / "mov.nii.gz")))
    fix = fileio.reorient_ari(fileio.read_nifti(str(tmp_path / "fix.nii")))
    rp = oracle.reg_params(symmetric=1, affine=affine, nw=7, ng=3, pyr_num=2, nthreads=8)
    ref_p, ref_c, _ = oracle.register(fix, mov, rp)
    M1, o1 = transforms.matrix_offset(params, c, affine)
    M2, o2 = transforms.matrix_offset(ref_p, ref_c, affine)
    dmm = transforms.max_corner_distance(M1, o1, M2, o2, pair.fixed)
    ddeg = transforms.rotation_angle_deg(M1, M2)
    print(f"cli vs oracle: corner distance {dmm:.3e} mm, rotation {ddeg:.3e} deg")
    assert dmm < 0.01 and ddeg < 0.01
    # --save-moving: the moving image resampled on the fixed grid under the
    # inverse transform with the default cubic B-spline.
    res = fileio.read_nifti(str(tmp_path / "res.nii.gz"))
    m_in = fileio.read_nifti(str(tmp_path / "mov.nii.gz"))
    f_in = fileio.read_nifti(str(tmp_path / "fix.nii"))
    assert res.dims == f_in.dims
    assert np.allclose(res.origin, f_in.origin) and np.allclose(res.spacing, f_in.spacing)
    Mi = np.linalg.inv(M1)
    ref = fr.resample_final(m_in.data, _G(m_in), Mi.ravel(), -Mi @ o1, _G(f_in), "bspline")
    bad = ~np.isclose(res.data, ref, rtol=1e-5, atol=1e-2)
    assert bad.mean() < 2e-3, (bad.sum(), np.abs(res.data - ref).max())


def test_nipype_interface_command_runs(tmp_path):
    """A registration driven through the restated nipype interface: its
    command line runs on the GPU build and every listed output exists."""
    from mirorr_amd.interface import Mirorr
    pair = phantom.make_pair(dims=(64, 64, 32), spacing=(3.0, 3.0, 3.0), n_ellipsoids=20,
                             seed=5, model="affine")
    fileio.write_nifti(str(tmp_path / "mov.nii.gz"), pair.moving)
    fileio.write_nifti(str(tmp_path / "fix.nii.gz"), pair.fixed)
    m = Mirorr(cwd=str(tmp_path), moving_image=str(tmp_path / "mov.nii.gz"),
               fixed_image=str(tmp_path / "fix.nii.gz"), transform_type="affine",
               pyramid_end=0, metric="NC", resampled_fixed="rf.nii.gz",
               resampled_moving="rm.nii.gz")
    print("interface", m.cmdline, flush=True)
    rc, out, err, outputs = m.run(timeout=180)
    assert rc == 0, err + out[-2000:]
    params, c = fileio.read_tfm(outputs["forward_transforms"][0], affine=True)
    assert np.all(np.isfinite(params)) and np.allclose(c, 0)
    for key, grid in (("resampled_moving", pair.fixed), ("resampled_fixed", pair.moving)):
        v = fileio.read_nifti(outputs[key])
        assert v.dims == grid.dims and np.count_nonzero(v.data) > 0.5 * v.data.size


def _crop_xy(v, cx, cy):
    """The volume without cx voxels at both ends of axis 0 and cy of axis 1."""
    d = np.ascontiguousarray(v.data[:, cy:v.data.shape[1] - cy, cx:v.data.shape[2] - cx])
    D = np.reshape(v.direction, (3, 3))
    org = np.asarray(v.origin) + D @ (np.asarray(v.spacing) * np.array([cx, cy, 0]))
    from mirorr_amd.phantom import Volume
    return Volume(d, v.spacing, tuple(org), v.direction)


def test_cli_crop_matches_oracle_on_cropped_volumes(tmp_path, oracle):
    pair = phantom.make_pair(dims=(64, 64, 32), spacing=(3.0, 3.0, 3.0), n_ellipsoids=20,
                             seed=11, model="rigid")
    fileio.write_nifti(str(tmp_path / "mov.nii.gz"), pair.moving)
    fileio.write_nifti(str(tmp_path / "fix.nii.gz"), pair.fixed)
    args = [CLI, "mov.nii.gz", "fix.nii.gz", "out.tfm", "-d", "2", "-c", "0", "--fresh",
            "--crop", "4,2,3,5"]
    print("cli", " ".join(args[1:]), flush=True)
    r = subprocess.run(args, cwd=tmp_path, capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stderr + r.stdout[-2000:]
    params, c = fileio.read_tfm(str(tmp_path / "out.tfm"), affine=False)
    mov = _crop_xy(fileio.reorient_ari(fileio.read_nifti(str(tmp_path / "mov.nii.gz"))), 4, 2)
    fix = _crop_xy(fileio.reorient_ari(fileio.read_nifti(str(tmp_path / "fix.nii.gz"))), 3, 5)
    rp = oracle.reg_params(symmetric=1, affine=0, nw=7, ng=3, pyr_num=2, nthreads=8)
    ref_p, ref_c, _ = oracle.register(fix, mov, rp)
    M1, o1 = transforms.matrix_offset(params, c, False)
    M2, o2 = transforms.matrix_offset(ref_p, ref_c, False)
    dmm = transforms.max_corner_distance(M1, o1, M2, o2, pair.fixed)
    print(f"crop: cli vs oracle {dmm:.3e} mm")
    assert dmm < 0.01


def test_cli_reorient_flags_give_the_same_registration(tmp_path):
    """Shifting both images to identity space before registering and
    composing the shifts back gives the registration of the original images
    (the shifts are translations of images already in ARI orientation)."""
    pair = phantom.make_pair(dims=(64, 64, 32), spacing=(3.0, 3.0, 3.0), n_ellipsoids=20,
                             seed=4, model="rigid")
    fileio.write_nifti(str(tmp_path / "mov.nii.gz"), pair.moving)
    fileio.write_nifti(str(tmp_path / "fix.nii.gz"), pair.fixed)
    out = {}
    for name, extra in (("plain", []), ("shifted", ["--reorient-moving", "--reorient-fixed"])):
        args = [CLI, "mov.nii.gz", "fix.nii.gz", f"{name}.tfm", "-d", "2", "-c", "0", "--fresh",
                *extra]
        print("cli", " ".join(args[1:]), flush=True)
        r = subprocess.run(args, cwd=tmp_path, capture_output=True, text=True, timeout=180)
        assert r.returncode == 0, r.stderr + r.stdout[-2000:]
        out[name] = transforms.matrix_offset(*fileio.read_tfm(str(tmp_path / f"{name}.tfm"),
                                                              affine=False), False)
    dmm = transforms.max_corner_distance(*out["plain"], *out["shifted"], pair.fixed)
    print(f"reorient flags: {dmm:.3e} mm")
    assert dmm < 0.05
