"""The lone-lane path, made to run on purpose.

A wave left with a single unfinished lane runs that lane in a loop of its own
with a branch-free type I lookup (lone_trips and cell2_flat in the HIP
sources).  Here it runs (a) from the first trip, on single-point calls, for
points on, near and outside the domain boundary, on square edges and on the
triangle diagonals, and (b) for ~1 000 trips, on the corner pixel (0, 0) of a
pixel grid (zero velocity there up to rounding: in the direction that drifts
it just outside the domain it falls back to the RK2 substeps at every step).
Points and fallback counts must equal the oracle's bit for bit
(cpab/gpu/calc_2Dto2D.cu:348-408 schedule, :38-153 lookup).
"""
import sys

import numpy as np
import pytest

from cpabdiffeo_amd import configs
from cpabdiffeo_amd.registration import pixel_basis

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


def dev(torch, x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def run_calc_T(torch, sp, T, sA, pts):
    out = torch.zeros(pts.shape, dtype=torch.float64, device="cuda")
    nfb = torch.zeros(len(pts), dtype=torch.int32, device="cuda")
    sp._gpu_calcs.calc_T(None, None, dev(torch, T), dev(torch, sA), dev(torch, pts), 0.01, 100, 10, out,
                         sp.dim_domain, sp.dim_range, sp.nCs, sp.incs, n_fallback=nfb)
    return out.cpu().numpy(), nfb.cpu().numpy()


@pytest.fixture(scope="module")
def pixel_field():
    n = 64
    su = configs.unit_space((4, 4), "I")
    sp = configs.pixel_space(n, (4, 4), "I")
    rng = np.random.default_rng(7)
    theta = 0.5 * su.sample_theta(su.se_prior_cov(0.1, 1.0), rng)
    Av = pixel_basis(su, n) @ theta
    return n, sp, Av.reshape(sp.nC, 6)


def test_single_points_on_and_outside_the_boundary(gpu, orc, pixel_field):
    torch = gpu
    n, sp, As = pixel_field
    g = sp.tessellation.geom()
    tiny = 1e-16
    special = [(0.0, 0.0), (n, 0.0), (0.0, n), (n, n), (-tiny, -tiny), (-tiny, 5.0), (5.0, -tiny),
               (n + 1e-9, n / 2), (n / 2, -3.0), (-2.0, n + 2.0), (16.0, 16.0), (16.0, 8.0), (10.0, 10.0),
               (10.0, 6.0), (n - 1e-11, n - 1e-11), (0.5, 0.5), (31.0, 33.0), (48.0, 0.0), (0.0, 48.0)]
    say("single-point calls:", len(special))
    for sign in (1, -1):
        sA, T = orc.prep(2, As, sign, 0.01)
        for xy in special:
            pts = np.array([xy], dtype=np.float64)
            want, nfb_w, _ = orc.calc_T(g, pts, T, sA, 0.01, 100, 10, probe=True)
            got, nfb = run_calc_T(torch, sp, T, sA, pts)
            assert np.array_equal(got, want), (sign, xy, got, want)
            assert np.array_equal(nfb, nfb_w), (sign, xy, nfb, nfb_w)


def test_corner_pixel_alone_for_a_thousand_trips(gpu, orc, pixel_field):
    torch = gpu
    n, sp, As = pixel_field
    g = sp.tessellation.geom()
    grid = sp.tessellation.dense_grid_img()
    corner = int(np.nonzero((grid == 0).all(1))[0][0])
    longest = 0
    for sign in (1, -1):  # the field leaves the corner one way: one sign drifts it outside
        sA, T = orc.prep(2, As, sign, 0.01)
        want, nfb_w, _ = orc.calc_T(g, grid, T, sA, 0.01, 100, 10, probe=True)
        say("dense grid", grid.shape, "sign", sign, "corner fallbacks", int(nfb_w[corner]), "max elsewhere",
            int(np.delete(nfb_w, corner).max()))
        longest = max(longest, int(nfb_w[corner]))
        got, nfb = run_calc_T(torch, sp, T, sA, grid)
        assert np.array_equal(got, want), sign
        assert np.array_equal(nfb, nfb_w), sign
    assert longest >= 50  # the corner ran long after its wave's other lanes
