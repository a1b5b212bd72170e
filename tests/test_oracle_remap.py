"""Pins of the oracle's remap and SSD likelihood.

The reference delegates remap to the un-vendored `of` package (parity
unpinned there); this project defines remap as OpenCV INTER_LINEAR with a
zero constant border, pinned here by:
* identity maps reproduce the image bit for bit;
* bilinear interpolation reproduces affine images exactly inside the grid;
* scipy.ndimage.map_coordinates(order=1) agrees on interior points;
* the half-pixel border blends with zero; far outside is zero.
The SSD likelihood is pinned by closed forms on constant images.
"""
import numpy as np
from scipy.ndimage import map_coordinates


def test_identity_map_copies(orc):
    rng = np.random.default_rng(0)
    img = rng.uniform(size=(33, 47)).astype(np.float32)
    yy, xx = np.mgrid[0:33, 0:47]
    pts = np.stack([xx.ravel(), yy.ravel()], 1).astype(np.float64)
    assert np.array_equal(orc.remap2d(pts, img), img)
    vol = rng.uniform(size=(5, 6, 7)).astype(np.float32)
    zz, yy, xx = np.mgrid[0:5, 0:6, 0:7]
    p3 = np.stack([xx.ravel(), yy.ravel(), zz.ravel()], 1).astype(np.float64)
    assert np.array_equal(orc.remap3d(p3, vol), vol)


def test_affine_image_exact_inside(orc):
    H, W = 40, 50
    yy, xx = np.mgrid[0:H, 0:W]
    img = (0.1 + 0.01 * xx + 0.02 * yy).astype(np.float32)
    rng = np.random.default_rng(1)
    pts = np.stack([rng.uniform(0, W - 1, H * W), rng.uniform(0, H - 1, H * W)], 1)
    out = orc.remap2d(pts, img).ravel()
    want = 0.1 + 0.01 * pts[:, 0] + 0.02 * pts[:, 1]
    np.testing.assert_allclose(out, want, rtol=2e-6, atol=2e-6)


def test_matches_map_coordinates_inside(orc):
    rng = np.random.default_rng(2)
    img = rng.uniform(size=(30, 30)).astype(np.float32)
    pts = rng.uniform(0, 29, (900, 2))
    want = map_coordinates(img.astype(np.float64), [pts[:, 1], pts[:, 0]], order=1)
    np.testing.assert_allclose(orc.remap2d(pts, img).ravel(), want, rtol=1e-5, atol=1e-6)


def test_border_blends_with_zero(orc):
    img = np.full((4, 4), 2.0, np.float32)
    pts = np.zeros((16, 2))
    pts[0] = (-0.5, 1.0)    # half a pixel left of column 0
    pts[1] = (3.5, 2.0)     # half a pixel right of the last column
    pts[2] = (-1.0, 1.0)    # a full pixel out: zero
    pts[3] = (1.0, -3.0)    # far above
    pts[4] = (np.nan, 1.0)
    pts[5:] = (1.0, 1.0)
    out = orc.remap2d(pts, img).ravel()
    assert out[0] == 1.0 and out[1] == 1.0 and out[2] == 0.0 and out[3] == 0.0 and out[4] == 0.0
    assert (out[5:] == 2.0).all()


def test_ssd_closed_form(orc):
    n = 1000
    x = np.full(n, 0.75, np.float32)
    y = np.full(n, 0.25, np.float32)
    assert orc.ssd_ll(x, y, 0.02) == -0.5 * n * 0.25 / 0.0004
    assert orc.ssd_ll(x, x, 0.02) == 0.0


def test_bilinear64_exact_on_affine_images_and_matches_scipy(orc):
    """The FP64 bilinear interpolant and gradient the SSD adjoint uses: exact
    on an affine image (values and gradient) away from the border, equal to
    scipy's order-1 map_coordinates inside, zero a full pixel outside."""
    import numpy as np
    from scipy.ndimage import map_coordinates
    H, W = 20, 30
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
    img = 0.5 + 0.03 * xx - 0.02 * yy
    rng = np.random.default_rng(4)
    pts = np.stack([rng.uniform(0, W - 1, 500), rng.uniform(0, H - 1, 500)], axis=1)
    w, gx, gy = orc.bilinear64(img, pts)
    np.testing.assert_allclose(w, 0.5 + 0.03 * pts[:, 0] - 0.02 * pts[:, 1], rtol=0, atol=1e-14)
    np.testing.assert_allclose(gx, 0.03, atol=1e-14)
    np.testing.assert_allclose(gy, -0.02, atol=1e-14)
    img2 = rng.random((H, W))
    w2, _, _ = orc.bilinear64(img2, pts)
    ref = map_coordinates(img2, [pts[:, 1], pts[:, 0]], order=1, mode="constant", cval=0.0)
    np.testing.assert_allclose(w2, ref, atol=1e-13)
    far = np.array([[-1.5, 3.0], [W + 0.5, 3.0], [3.0, H], [np.nan, 1.0]])
    w3, g3x, g3y = orc.bilinear64(img2, far)
    assert (w3 == 0).all() and (g3x == 0).all() and (g3y == 0).all()
    # gradient = central differences of the interpolant away from pixel edges
    q = np.array([[3.3, 4.6], [10.25, 7.75]])
    e = 1e-7
    for k in range(2):
        d = np.zeros(2); d[k] = e
        num = (orc.bilinear64(img2, q + d)[0] - orc.bilinear64(img2, q - d)[0]) / (2 * e)
        np.testing.assert_allclose(orc.bilinear64(img2, q)[1 + k], num, rtol=1e-7)


def test_trilinear_exact_on_affine_volumes_at_fractional_coordinates(orc):
    """Trilinear remap at fractional coordinates (not only the identity map):
    exact on an affine volume with a different slope per axis (a swapped or
    transposed axis fails), and equal to scipy's order-1 map_coordinates on a
    random volume inside."""
    D, H, W = 12, 17, 23
    zz, yy, xx = np.mgrid[0:D, 0:H, 0:W]
    vol = (0.25 + 0.015625 * xx + 0.03125 * yy - 0.0625 * zz).astype(np.float32)  # exact in float32
    rng = np.random.default_rng(5)
    n = D * H * W
    pts = np.stack([rng.uniform(0, W - 1, n), rng.uniform(0, H - 1, n), rng.uniform(0, D - 1, n)], 1)
    out = orc.remap3d(pts, vol).ravel()
    want = 0.25 + 0.015625 * pts[:, 0] + 0.03125 * pts[:, 1] - 0.0625 * pts[:, 2]
    np.testing.assert_allclose(out, want, rtol=0, atol=4e-6)
    v2 = rng.uniform(size=(D, H, W)).astype(np.float32)
    ref = map_coordinates(v2.astype(np.float64), [pts[:, 2], pts[:, 1], pts[:, 0]], order=1)
    np.testing.assert_allclose(orc.remap3d(pts, v2).ravel(), ref, rtol=1e-5, atol=2e-6)
