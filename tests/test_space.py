"""Host-side geometry: tessellation order and increments, the CPA basis, the
prior, and the configuration generators (no GPU needed)."""
import numpy as np
import pytest

from cpabdiffeo_amd import configs
from cpabdiffeo_amd.space import CpaSpace
from cpabdiffeo_amd.tessellation import Tessellation


def test_increments_match_the_reference_formula():
    assert Tessellation([0], [1], [50], "II").incs[0] == 0.019999999999999997
    assert list(Tessellation([0, 0], [1, 1], [4, 4], "I").incs) == [0.25, 0.25]
    assert list(Tessellation([0, 0], [512, 512], [4, 4], "I").incs) == [128.0, 128.0]
    assert list(Tessellation([0, 0, 0], [64, 64, 64], [4, 4, 4], "I").incs) == [16.0] * 3
    # one cell along an axis: twice the only centre
    assert list(Tessellation([0, 0], [3, 5], [1, 1], "II").incs) == [3.0, 5.0]


def test_type_I_triangle_order():
    T = Tessellation([0, 0], [2, 2], [2, 2], "I")
    v = T.cells_verts_homo_coo
    # rect (0,0): centre (0.5, 0.5); triangles top, right, bottom, left
    assert np.array_equal(v[0, :, :2], [[0.5, 0.5], [0, 0], [1, 0]])
    assert np.array_equal(v[1, :, :2], [[0.5, 0.5], [1, 0], [1, 1]])
    assert np.array_equal(v[2, :, :2], [[0.5, 0.5], [1, 1], [0, 1]])
    assert np.array_equal(v[3, :, :2], [[0.5, 0.5], [0, 1], [0, 0]])
    # next rectangle along x
    assert np.array_equal(v[4, 0, :2], [1.5, 0.5])


@pytest.mark.parametrize("args,d", [
    (([0, 0], [1, 1], [4, 4], [True, True], "I", False), 62),
    (([0], [1], [50], [True], "II", None), 49),
    (([0, 0, 0], [1, 1, 1], [4, 4, 4], [True] * 3, "I", False), 225),
])
def test_basis_dimension_and_constraints(args, d):
    sp = CpaSpace(args[0], args[1], args[2], zero_v_across_bdry=args[3], tess=args[4],
                  valid_outside=args[5])
    assert sp.d == d
    assert np.abs(sp.B.T @ sp.B - np.eye(d)).max() < 1e-12
    assert np.abs(sp.constraintMat @ sp.B).max() < 1e-12


def test_fields_are_continuous_and_tangential_on_the_boundary():
    sp = configs.unit_space((4, 4), "I")
    rng = np.random.default_rng(0)
    Av = sp.B @ rng.normal(size=sp.d)
    A = Av.reshape(sp.nC, 2, 3)
    X = sp.tessellation.cells_verts_homo_coo
    vals = {}
    for c in range(sp.nC):
        for v in X[c]:
            vel = A[c] @ v
            key = tuple(v[:2])
            if key in vals:
                assert np.abs(vals[key] - vel).max() < 1e-12  # continuity at shared vertices
            vals[key] = vel
            if v[0] in (0.0, 1.0):
                assert abs(vel[0]) < 1e-12  # zero normal velocity
            if v[1] in (0.0, 1.0):
                assert abs(vel[1]) < 1e-12


def test_theta_avees_roundtrip_and_veltess():
    sp = configs.unit_space((4, 4), "I")
    rng = np.random.default_rng(1)
    theta = rng.normal(size=sp.d)
    Av = sp.theta2Avees(theta).copy()
    assert np.allclose(sp.Avees2theta(Av), theta)
    vt = sp.Avees2velTess(Av)
    assert np.allclose(sp.velTess2Avees(vt, np.zeros(sp.D)), Av)


def test_prior_is_spd_and_seeded():
    cfg = configs.config1(grid=8)
    cov = cfg["cov"]
    assert np.allclose(cov, cov.T)
    assert np.linalg.eigvalsh(cov).min() > -1e-10
    again = configs.config1(grid=8)
    assert np.array_equal(cfg["Avees"], again["Avees"])


def test_unit_to_pixel_is_the_same_field():
    sp_u = configs.unit_space((4, 4), "I")
    sp_p = configs.pixel_space(256)
    rng = np.random.default_rng(2)
    Av = sp_u.B @ rng.normal(size=sp_u.d)
    Ap = configs.unit_to_pixel_avees(sp_u, Av, 256).reshape(sp_p.nC, 2, 3)
    Au = Av.reshape(sp_u.nC, 2, 3)
    x = np.array([0.3, 0.7, 1.0])
    X = np.array([0.3 * 256, 0.7 * 256, 1.0])
    assert np.allclose(Ap[5] @ X, 256 * (Au[5] @ x))


def test_blob_image():
    img = configs.blob_image(64, seed=3)
    assert img.dtype == np.float32 and img.shape == (64, 64)
    assert img.min() == 0.0 and img.max() == 1.0
