"""Pins the oracle's element matrices (R/src/Core/Element.jl:8-134) against the
closed-form golden vectors and against what the mathematics fixes."""
import os

import numpy as np
import pytest

import oracle

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "golden.npz"))


@pytest.mark.parametrize("tag", ["unit", "aniso", "h2nu25"])
def test_ke_matches_closed_form(tag):
    p = GOLD[f"ke_{tag}_params"]
    K = oracle.ke_hex8(*p[:3], nu=p[3])
    ref = GOLD[f"ke_{tag}"]
    assert np.abs(K - ref).max() <= 1e-14 * np.abs(ref).max()


@pytest.mark.parametrize("tag", ["unit", "aniso", "h2nu25"])
def test_ke_f32_widened_is_float_rounding_of_exact(tag):
    p = GOLD[f"ke_{tag}_params"]
    Kf = oracle.ke_hex8(*p[:3], nu=p[3], f32=True)
    ref = GOLD[f"ke_{tag}"]
    err = np.abs(Kf - ref).max() / np.abs(ref).max()
    assert 0 < err < 2e-6          # float evaluation: differs, but only at float rounding
    assert np.all(Kf.astype(np.float32).astype(np.float64) == Kf)


def rigid_modes():
    X = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    modes = []
    for c in range(3):
        t = np.zeros((8, 3)); t[:, c] = 1; modes.append(t.ravel())
    for a, b in ((0, 1), (1, 2), (0, 2)):
        r = np.zeros((8, 3)); r[:, a] = -X[:, b]; r[:, b] = X[:, a]; modes.append(r.ravel())
    return np.array(modes).T


def test_ke_invariants():
    """Symmetric, PSD, exactly 6 rigid-body zero modes, patch test energy."""
    nu = 0.3
    K = oracle.ke_hex8(1.0, 1.0, 1.0, nu)
    assert np.allclose(K, K.T, atol=1e-15)
    w = np.linalg.eigvalsh(K)
    assert np.all(w[:6] < 1e-13) and w[6] > 1e-2
    R = rigid_modes()
    assert np.abs(K @ R).max() < 1e-13
    # uniform strain eps_xx = 1: u_x = x -> energy u'Ku = V eps' D eps = c1
    X = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    u = np.zeros((8, 3)); u[:, 0] = X[:, 0]
    c1 = (1 - nu) / ((1 + nu) * (1 - 2 * nu))
    assert abs(u.ravel() @ K @ u.ravel() - c1) < 1e-13
    # pure shear gamma_xy = 1: u_x = y -> energy = mu
    u = np.zeros((8, 3)); u[:, 0] = X[:, 1]
    assert abs(u.ravel() @ K @ u.ravel() - 1 / (2 * (1 + nu))) < 1e-13


def test_ke_f32_widened_breaks_the_rigid_body_null_space():
    """The Float32-evaluated Ke (the reference's element matrix) no longer
    annihilates rigid translations, and its symmetric part is indefinite on
    the rigid-body subspace: the reason the exact Ke is the drop-in default."""
    Kw = oracle.ke_hex8(1.0, 1.0, 1.0, 0.3, f32=True)
    K = oracle.ke_hex8(1.0, 1.0, 1.0, 0.3)
    R = rigid_modes()
    assert np.abs(K @ R[:, :3]).max() < 1e-15
    assert 1e-9 < np.abs(Kw @ R[:, :3]).max() < 1e-6
    w = np.linalg.eigvalsh(0.5 * (Kw + Kw.T))
    assert w[0] < -1e-9 and abs(w[0]) < 1e-7 and w[6] > 1e-2


def test_ke_scales_with_size():
    """Ke(a h) = a Ke(h) for 3-D elasticity (used by the multigrid levels)."""
    K1 = oracle.ke_hex8(1.0, 2.0, 0.5, 0.3)
    K2 = oracle.ke_hex8(2.0, 4.0, 1.0, 0.3)
    assert np.abs(K2 - 2 * K1).max() < 1e-13


def test_ke_rejects_bad_input():
    with pytest.raises(ValueError):
        oracle.ke_hex8(1.0, 1.0, 1.0, 0.5)
    with pytest.raises(ValueError):
        oracle.ke_hex8(0.0, 1.0, 1.0, 0.3)


@pytest.mark.parametrize("bug", [True, False])
def test_helmholtz_matrices(bug):
    K8, M8 = oracle.helm_mats(bug_compat=bug)
    tag = "bug" if bug else "corr"
    assert np.abs(K8 - GOLD[f"helm_K8_{tag}"]).max() < 1e-15
    assert np.abs(M8 - GOLD[f"helm_M8_{tag}"]).max() < 1e-15
    # mass sums to the element volume: 1.5 V with the reference's z=dz node table
    assert abs(M8.sum() - (1.5 if bug else 1.0)) < 1e-14
    # Laplacian annihilates constants
    assert np.abs(K8.sum(axis=1)).max() < 1e-14
