"""The order in which the kernels take their points changes no result.

`cpab_calc_T` guesses a row-major grid's row length from the points and then
takes them in 8 x 8 tiles (the rows past the last whole band of 8 in index
order); the batch sorts each tile's points by position in LDS.  Each point's
end point and fallback count must still be the oracle's bit for bit, for
grids with partial bands, grids too short to tile, false guesses (points
that only start like a grid), column-major grids and random points, and for
batch tiles of every size around the 16 384-point limit of the sort.
"""
import sys

import numpy as np
import pytest

from cpabdiffeo_amd import configs

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


def _grid(W, H, order="row"):
    """The pixel-centre-free grid x/W, y/H of the unit square, row-major
    (x fastest) or column-major."""
    yy, xx = np.mgrid[0:H, 0:W]
    if order == "col":
        yy, xx = yy.T, xx.T
    return np.stack([xx.ravel() / W, yy.ravel() / H], 1).astype(np.float64)


def _cases():
    rng = np.random.default_rng(7)
    c = {
        "row 64x64 (whole bands)": _grid(64, 64),
        "row 40x37 (4 bands + 5 rows in index order)": _grid(40, 37),
        "row 24x5 (fewer than 8 rows: no tiles)": _grid(24, 5),
        "row 8x100 (one tile a band)": _grid(8, 100),
        "row 2048x9 (the longest row a block finds)": _grid(2048, 9),
        "col 48x48 (column-major: no row length)": _grid(48, 48, "col"),
        "random": rng.uniform(0, 1, (5000, 2)),
    }
    fake = rng.uniform(0, 1, (3000, 2))
    fake[:16, 1] = 0.25  # y constant over the first 16 points, then anything
    c["false guess (16 equal y, then random)"] = fake
    return c


CASES = _cases()


@pytest.mark.parametrize("name", list(CASES))
def test_calc_T_any_point_order_is_bitexact(gpu, orc, name):
    torch = gpu
    pts = CASES[name]
    say("calc_T order case:", name, len(pts), "points")
    sp = configs.unit_space((4, 4), "I")
    g = sp.tessellation.geom()
    rng = np.random.default_rng(11)
    Av = sp.B @ sp.sample_theta(sp.se_prior_cov(0.1, 1.0), rng)
    sA, T = orc.prep(2, Av.reshape(sp.nC, -1), 1, 0.01)
    want, nfb_w, _ = orc.calc_T(g, pts, T, sA, 0.01, 100, 10, probe=True)
    out = torch.zeros(pts.shape, dtype=torch.float64, device="cuda")
    nfb = torch.zeros(len(pts), dtype=torch.int32, device="cuda")
    d = lambda x: torch.from_numpy(np.ascontiguousarray(x)).cuda()
    sp._gpu_calcs.calc_T(None, None, d(T), d(sA), d(pts), 0.01, 100, 10, out, 2, 2, sp.nCs, sp.incs, n_fallback=nfb)
    got = out.cpu().numpy()
    bad = np.nonzero((got != want).any(1))[0]
    assert bad.size == 0, (name, bad[:5])
    assert np.array_equal(nfb.cpu().numpy(), nfb_w)


@pytest.mark.parametrize("tile", [1024, 16384, 65536])
def test_batch_tiles_around_the_sort_limit(gpu, orc, tile):
    """Instances of 1, 7, 16 383, 16 384 and 16 385 points (the last splits
    into a full tile and a one-point tile), uniform random and on a grid, at
    tile sizes below, at and above the sorted-tile limit: the oracle's bits."""
    torch = gpu
    from cpabdiffeo_amd import batch
    say("batch tiles of", tile)
    rng = np.random.default_rng(tile)
    sizes = [1, 7, 16383, 16384, 16385]
    spaces = {4: configs.unit_space((4, 4), "I"), 8: configs.unit_space((8, 8), "I")}
    insts, As, pts = [], [], []
    for k, n in enumerate(sizes):
        grid = 4 if k % 2 == 0 else 8
        sp = spaces[grid]
        Av = sp.B @ sp.sample_theta(sp.se_prior_cov(0.1, 1.0), rng)
        insts.append((grid, Av, n))
        As.append(Av.reshape(sp.nC, -1))
        p = rng.uniform(0, 1, (n, 2)) if k != 3 else _grid(128, 128)
        pts.append(p)
    plan = batch.BatchPlan(2, "I", [((g, g), (1.0 / g, 1.0 / g), n) for g, _, n in insts])
    allp = np.concatenate(pts)
    d = lambda x: torch.from_numpy(np.ascontiguousarray(x)).cuda()
    tables = plan.prep(d(np.concatenate(As)), 1, 0.01)
    out = torch.empty((plan.nPts, 2), dtype=torch.float64, device="cuda")
    nfb = torch.zeros(plan.nPts, dtype=torch.int32, device="cuda")
    plan.calc_T(tables, d(allp), out, 0.01, 100, 10, sched=batch.BatchSched(tile, 1024, 1), use_graph=False,
                n_fallback=nfb)
    got, gotn = out.cpu().numpy(), nfb.cpu().numpy()
    tab = tables.cpu().numpy()
    off = 0
    for i, (grid, _, n) in enumerate(insts):
        sp = spaces[grid]
        t = tab[off:off + 2 * sp.nC * 6].reshape(2, sp.nC, 6)
        off += 2 * sp.nC * 6
        o0, o1 = plan.offsets[i], plan.offsets[i + 1]
        want, nfb_w, _ = orc.calc_T(sp.tessellation.geom(), allp[o0:o1], t[0], t[1], 0.01, 100, 10, probe=True)
        assert np.array_equal(got[o0:o1], want), (tile, i, n)
        assert np.array_equal(gotn[o0:o1], nfb_w), (tile, i, n)
