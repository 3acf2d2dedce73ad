"""Pins the oracle's LTS regression and transform algebra
(itkMirorrRegistrationMethod.txx:758-1106, itkSymMirorrRegistrationMethod.txx:729-1442,
vnl_utils.h:46-135) against closed forms and numpy/scipy library routines."""
import numpy as np
import pytest
import scipy.linalg
from scipy.spatial.transform import Rotation


def _rot(ax, ay, az):
    """Rotation R = Rz Ry Rx (the convention of vnl's Euler angles)."""
    return Rotation.from_euler("xyz", [ax, ay, az]).as_matrix()


def _kabsch(X, Y):
    """Least-squares rotation mapping centred X onto centred Y (numpy SVD)."""
    H = X.T @ Y
    U, _, Vt = np.linalg.svd(H)
    d = np.sign(np.linalg.det(Vt.T @ U.T))
    D = np.diag([1, 1, d])
    return Vt.T @ D @ U.T


def _quat_matrix(q):
    x, y, z, r = q
    return Rotation.from_quat([x, y, z, r]).as_matrix()


def test_euler3d_matrix_is_zxy(oracle):
    ang = np.array([0.2, -0.3, 0.5])
    rx = Rotation.from_euler("x", ang[0]).as_matrix()
    ry = Rotation.from_euler("y", ang[1]).as_matrix()
    rz = Rotation.from_euler("z", ang[2]).as_matrix()
    np.testing.assert_allclose(oracle.euler_matrix(ang), rz @ rx @ ry, atol=1e-15)


def test_quaternion_euler_angles_are_zyx(oracle):
    q = Rotation.from_euler("xyz", [0.3, -0.2, 0.7]).as_quat()  # (x, y, z, w)
    np.testing.assert_allclose(oracle.quat_euler(q), [0.3, -0.2, 0.7], atol=1e-12)
    # q and -q give the same angles
    np.testing.assert_allclose(oracle.quat_euler(-q), [0.3, -0.2, 0.7], atol=1e-12)


def test_rigid_quaternion_equals_kabsch(oracle):
    rng = np.random.default_rng(0)
    X = rng.normal(0, 30, (200, 3))
    R = _rot(0.1, -0.15, 0.2)
    Y = X @ R.T + rng.normal(0, 0.5, X.shape)
    Xc, Yc = X - X.mean(0), Y - Y.mean(0)
    q = oracle.rigid_quaternion(Xc, Yc)
    np.testing.assert_allclose(np.linalg.norm(q), 1.0, atol=1e-12)
    np.testing.assert_allclose(_quat_matrix(q), _kabsch(Xc, Yc), atol=1e-9)


def test_orthogonalize_is_polar_factor(oracle):
    rng = np.random.default_rng(1)
    m = np.eye(3) + rng.normal(0, 0.1, (3, 3))
    U, _, Vt = np.linalg.svd(m)
    np.testing.assert_allclose(oracle.orthogonalize(m), U @ Vt, atol=1e-12)


def test_denman_beavers_sqrt(oracle):
    A = np.eye(4)
    A[:3, :3] = _rot(0.2, 0.1, -0.3) @ np.diag([1.1, 0.95, 1.02])
    A[:3, 3] = [5.0, -3.0, 2.0]
    S = oracle.sqrtm4(A)
    np.testing.assert_allclose(S @ S, A, atol=1e-12)
    np.testing.assert_allclose(S, np.real(scipy.linalg.sqrtm(A)), atol=1e-10)


def test_classic_lts_pure_translation_exact(oracle):
    rng = np.random.default_rng(2)
    X = rng.uniform(-60, 60, (300, 3))
    t = np.array([3.5, -2.0, 7.25])
    Y = X + t
    for kind in (0, 1):
        n = 6 if kind == 0 else 12
        p0 = np.zeros(n)
        if kind == 1:
            p0[[0, 4, 8]] = 1.0
        p, c, it = oracle.lts_classic(kind, X, Y, p0, [1.0, 2.0, 3.0])
        np.testing.assert_allclose(p[n - 3:], t, atol=1e-10)
        assert it >= 1


def test_symmetric_affine_exact_and_inverse_consistent(oracle):
    rng = np.random.default_rng(3)
    X = rng.uniform(-80, 80, (400, 3))
    A = _rot(0.1, -0.05, 0.12) @ np.array([[1.05, 0.02, 0.0], [0.0, 0.97, 0.03], [0.0, 0.0, 1.08]])
    t = np.array([4.0, -6.0, 2.5])
    Y = X @ A.T + t
    f, b, p, _ = oracle.lts_symmetric(1, X, Y)
    F = np.vstack([f, [0, 0, 0, 1]])
    B = np.vstack([b, [0, 0, 0, 1]])
    T = F @ np.linalg.inv(B)  # final transform = t12F o t12B^-1
    np.testing.assert_allclose(T[:3, :3], A, atol=1e-9)
    np.testing.assert_allclose(T[:3, 3], t, atol=1e-8)
    # swapping the point sets inverts the fitted transform
    f2, b2, _, _ = oracle.lts_symmetric(1, Y, X)
    T2 = np.vstack([f2, [0, 0, 0, 1]]) @ np.linalg.inv(np.vstack([b2, [0, 0, 0, 1]]))
    np.testing.assert_allclose(T2, np.linalg.inv(T), atol=1e-8)


def test_symmetric_affine_trims_gross_outliers(oracle):
    rng = np.random.default_rng(4)
    X = rng.uniform(-80, 80, (500, 3))
    A = np.diag([1.02, 0.98, 1.01])
    t = np.array([1.0, 2.0, -3.0])
    Y = X @ A.T + t
    bad = rng.choice(500, 150, replace=False)
    Y[bad] += rng.uniform(40, 80, (150, 3)) * rng.choice([-1, 1], (150, 3))
    f, b, _, _ = oracle.lts_symmetric(1, X, Y)
    T = np.vstack([f, [0, 0, 0, 1]]) @ np.linalg.inv(np.vstack([b, [0, 0, 0, 1]]))
    np.testing.assert_allclose(T[:3, :3], A, atol=1e-8)
    np.testing.assert_allclose(T[:3, 3], t, atol=1e-7)


def test_symmetric_rigid_single_axis_exact(oracle):
    rng = np.random.default_rng(5)
    X = rng.uniform(-80, 80, (300, 3))
    R = Rotation.from_euler("z", 0.15).as_matrix()
    t = np.array([3.0, -1.0, 2.0])
    Y = X @ R.T + t
    f, b, p, _ = oracle.lts_symmetric(0, X, Y)
    T = np.vstack([f, [0, 0, 0, 1]]) @ np.linalg.inv(np.vstack([b, [0, 0, 0, 1]]))
    np.testing.assert_allclose(T[:3, :3], R, atol=1e-9)
    np.testing.assert_allclose(T[:3, 3], t, atol=1e-8)
    # the half transforms are square roots of each other's inverse
    np.testing.assert_allclose(np.vstack([f, [0, 0, 0, 1]]) @ np.vstack([b, [0, 0, 0, 1]]),
                               np.eye(4), atol=1e-9)


def test_symmetric_rigid_general_rotation_is_swap_symmetric(oracle):
    """For a general rotation the Euler angles of vnl (Z*Y*X) feed ITK's
    Euler3D (Z*X*Y) and the fit is not exact, but it stays inverse-consistent."""
    rng = np.random.default_rng(6)
    X = rng.uniform(-80, 80, (300, 3))
    R = _rot(0.12, -0.08, 0.1)
    Y = X @ R.T + np.array([2.0, 1.0, -4.0])
    f, b, _, _ = oracle.lts_symmetric(0, X, Y)
    f2, b2, _, _ = oracle.lts_symmetric(0, Y, X)
    np.testing.assert_allclose(f2, b, atol=1e-9)
    np.testing.assert_allclose(b2, f, atol=1e-9)
    T = np.vstack([f, [0, 0, 0, 1]]) @ np.linalg.inv(np.vstack([b, [0, 0, 0, 1]]))
    assert np.allclose(T[:3, :3] @ T[:3, :3].T, np.eye(3), atol=1e-9)
