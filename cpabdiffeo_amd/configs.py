"""Synthetic workloads of the benchmark configurations, generated from seeds.

Each function returns plain numpy data (Avees tables, points, images); the
reference has no stored inputs for this path, so these generators define the
inputs that the GPU path, the CPU oracle and the fixtures share.

* ``config1``: [0,1]^2, 4x4 rectangles x 4 triangles (nC = 64), zero normal
  velocity on the boundary, theta ~ N(0, Sigma_CPA) with a squared-
  exponential kernel (length scale 0.1, variance 1) over the cell centroids,
  seed 0; t = 1, nTimeSteps = 100, nStepsODE = 10; points on a 512 x 512 grid
  (x = j/512, y = i/512, row-major, x fastest).
* ``blob_image``: sum of isotropic Gaussian blobs, min-max normalised, float32.
* ``config2``: 256 x 256 registration problem on the config-1 prior: the
  target is the source remapped by T^{-1} of a prior draw plus N(0, 0.02^2)
  pixel noise.  Points and images live in pixel units [0, 256]^2.
"""
import numpy as np

from .space import CpaSpace

CFG1 = dict(nCs=(4, 4), tess="I", t=1.0, nTimeSteps=100, nStepsODE=10, length_scale=0.1,
            variance=1.0, grid=512, seed=0)


def unit_space(nCs=(4, 4), tess="I", dim=2, zero_bdry=True):
    """CPA space on [0,1]^dim with zero normal velocity on the boundary."""
    if dim == 1:
        return CpaSpace([0], [1], list(nCs), zero_v_across_bdry=[zero_bdry], tess="II")
    valid_outside = not zero_bdry
    return CpaSpace([0] * dim, [1] * dim, list(nCs), zero_v_across_bdry=[zero_bdry] * dim,
                    tess=tess, valid_outside=valid_outside)


def grid_points(n, extent=1.0):
    """n x n grid (x = j*extent/n, y = i*extent/n), row-major with x fastest."""
    yy, xx = np.mgrid[0:n, 0:n]
    return (np.stack([xx.ravel(), yy.ravel()], axis=1) * (extent / n)).astype(np.float64)


def prior_avees(space, seed, length_scale=0.1, variance=1.0, n=None):
    """Avees = B theta with theta ~ N(0, Sigma_CPA) drawn from numpy's PCG64(seed)."""
    cov = space.se_prior_cov(length_scale=length_scale, variance=variance)
    rng = np.random.default_rng(seed)
    theta = space.sample_theta(cov, rng, n)
    if n is None:
        return space.B @ theta, theta, cov
    return theta @ space.B.T, theta, cov


def config1(grid=512):
    sp = unit_space(CFG1["nCs"], CFG1["tess"])
    Avees, theta, cov = prior_avees(sp, CFG1["seed"], CFG1["length_scale"], CFG1["variance"])
    pts = grid_points(grid)
    dt = CFG1["t"] / CFG1["nTimeSteps"]
    return dict(space=sp, Avees=Avees, theta=theta, cov=cov, pts=pts, dt=dt,
                nTimeSteps=CFG1["nTimeSteps"], nStepsODE=CFG1["nStepsODE"])


def blob_image(n, n_blobs=40, seed=0, dims=2):
    """Sum of isotropic Gaussian blobs (centres U[0,1]^d, sigma U[0.02,0.1],
    amplitude U[0.2,1]) on an n^d pixel grid, min-max normalised to [0,1]."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(0, 1, (n_blobs, dims))
    s = rng.uniform(0.02, 0.1, n_blobs)
    a = rng.uniform(0.2, 1.0, n_blobs)
    axes = [(np.arange(n) + 0.5) / n] * dims
    grids = np.meshgrid(*axes, indexing="ij")  # grids[0] = slowest axis
    img = np.zeros((n,) * dims)
    for k in range(n_blobs):
        d2 = sum((grids[dims - 1 - i] - c[k, i]) ** 2 for i in range(dims))
        img += a[k] * np.exp(-d2 / (2 * s[k] ** 2))
    img = (img - img.min()) / (img.max() - img.min())
    return img.astype(np.float32)


def pixel_space(n, nCs=(4, 4), tess="I"):
    """The config-1 tessellation on the pixel domain [0, n]^2."""
    return CpaSpace([0, 0], [n, n], list(nCs), zero_v_across_bdry=[True, True], tess=tess,
                    valid_outside=False)


def unit_to_pixel_avees(space_unit, Avees_unit, n):
    """Rescale a unit-domain field to [0, n]^2: the linear part is kept and
    the translation column scaled by n (same velocity field in pixel units,
    time unchanged)."""
    A = np.asarray(Avees_unit).reshape(space_unit.nC, space_unit.dim_domain, -1).copy()
    A[:, :, -1] *= n
    return A.reshape(-1)
