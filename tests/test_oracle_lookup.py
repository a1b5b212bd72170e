"""Pins of the oracle's cell lookup (calc_2Dto2D.cu:38-153, calc_1Dto1D.cu:62-65,
calc_3Dto3D.cu:43-137) against geometry computed independently:

* interior points: exact rational point-in-simplex tests against the cells of
  cpabdiffeo_amd.Tessellation (which also pins the cell ORDER the lookup
  assumes: the reference's Tessellation order);
* type II / 1D: floor of the coordinate in cell units, clamped;
* the out-of-domain extension rules of the 2D triangulation: hand-derived
  cases (each outside region maps to the boundary triangle facing it; the
  corner regions split along the diagonal through the corner).
"""
from fractions import Fraction

import numpy as np
import pytest

from cpabdiffeo_amd.tessellation import Tessellation


def bary_inside(verts, p):
    """Exact barycentric test: is p inside the simplex (strictly, by > eps)?"""
    n = len(p)
    V = [[Fraction(x) for x in v[:n]] for v in verts]
    P = [Fraction(x) for x in p]
    # solve sum_i l_i V_i = P, sum l_i = 1 with Fractions (Gaussian elimination)
    M = [[V[j][i] for j in range(n + 1)] + [P[i]] for i in range(n)] + [[Fraction(1)] * (n + 1) + [Fraction(1)]]
    m = n + 1
    for c in range(m):
        piv = next(r for r in range(c, m) if M[r][c] != 0)
        M[c], M[piv] = M[piv], M[c]
        for r in range(m):
            if r != c and M[r][c] != 0:
                f = M[r][c] / M[c][c]
                M[r] = [a - f * b for a, b in zip(M[r], M[c])]
    lam = [M[i][m] / M[i][i] for i in range(m)]
    return min(lam)


@pytest.mark.parametrize("nCs,extent", [((4, 4), 1.0), ((3, 5), 1.0), ((4, 4), 512.0), ((2, 3), 7.0)])
def test_2d_type_I_interior_point_in_triangle(orc, nCs, extent):
    T = Tessellation([0, 0], [extent, extent], nCs, "I")
    rng = np.random.default_rng(10)
    pts = rng.uniform(0, extent, (3000, 2))
    cells = orc.calc_cell_idx(T.geom(), pts)
    ok = 0
    for p, c in zip(pts, cells):
        lam = bary_inside(T.cells_verts_homo_coo[c], p)
        if lam > 1e-9:
            ok += 1
        else:
            # the point must then be within 1e-9 of a facet of its cell
            assert lam > -1e-9, (p, c, float(lam))
    assert ok > 2900


@pytest.mark.parametrize("nCs,extents", [((2, 2, 2), (1.0,) * 3), ((4, 4, 4), (64.0,) * 3),
                                          ((3, 2, 2), (1.5, 1.0, 1.0))])
def test_3d_type_I_interior_point_in_tetra(orc, nCs, extents):
    # cubic cells (inc_x == inc_z): the z clamp quirk below is then harmless
    T = Tessellation([0, 0, 0], list(extents), nCs, "I")
    extent = np.asarray(extents)
    rng = np.random.default_rng(11)
    pts = rng.uniform(0, 1, (1500, 3)) * extent
    cells = orc.calc_cell_idx(T.geom(), pts)
    ok = 0
    for p, c in zip(pts, cells):
        lam = bary_inside(T.cells_verts_homo_coo[c], p)
        if lam > 1e-9:
            ok += 1
        else:
            assert lam > -1e-9, (p, c, float(lam))
    assert ok > 1450


@pytest.mark.parametrize("dim,nCs", [(1, (50,)), (2, (4, 3)), (3, (2, 3, 4))])
def test_type_II_floor(orc, dim, nCs):
    T = Tessellation([0] * dim, [1.0] * dim, nCs, "II")
    rng = np.random.default_rng(12)
    pts = rng.uniform(-0.3, 1.3, (5000, dim))
    cells = orc.calc_cell_idx(T.geom(), pts)
    idx = [np.clip(np.floor(pts[:, d] * nCs[d]).astype(int), 0, nCs[d] - 1) for d in range(dim)]
    want = idx[0].copy()
    mult = 1
    for d in range(1, dim):
        mult *= nCs[d - 1]
        want += idx[d] * mult
    # points within an ulp of a cell wall may round either way: exclude them
    frac = np.abs(pts * np.asarray(nCs) - np.round(pts * np.asarray(nCs)))
    clear = (frac > 1e-9).all(axis=1)
    assert clear.sum() > 4900
    assert np.array_equal(cells[clear], want[clear])


def test_1d_fifty_cells_increment(orc):
    T = Tessellation([0], [1], [50], "II")
    assert T.incs[0] == 0.019999999999999997
    pts = (np.arange(50) + 0.5)[:, None] / 50
    assert np.array_equal(orc.calc_cell_idx(T.geom(), pts), np.arange(50))


def test_2d_type_I_outside_rules(orc):
    # 4 x 4 rectangles on [0,1]^2, inc = 0.25; rect (i,j) has triangles
    # 4*(i + 4j) + {0 top, 1 right, 2 bottom, 3 left}
    T = Tessellation([0, 0], [1, 1], (4, 4), "I")
    cases = [
        ((-0.10, 0.30), 4 * (0 + 1 * 4) + 3),   # left of row 1: left triangle
        ((1.20, 0.60), 4 * (3 + 2 * 4) + 1),    # right of row 2: right triangle
        ((0.60, -0.20), 4 * (2 + 0 * 4) + 0),   # above column 2: top triangle
        ((0.30, 1.10), 4 * (1 + 3 * 4) + 2),    # below column 1: bottom triangle
        ((-0.10, -0.30), 0),                    # top-left corner, steeper than the diagonal: top
        ((-0.30, -0.10), 3),                    # top-left corner, flatter: left
        ((-0.10, 1.30), 4 * (0 + 3 * 4) + 2),   # bottom-left corner, below the diagonal: bottom
        ((-0.30, 1.10), 4 * (0 + 3 * 4) + 3),   # bottom-left corner: left
        ((1.10, -0.30), 4 * 3 + 0),             # top-right corner: top
        ((1.30, -0.10), 4 * 3 + 1),             # top-right corner: right
        ((1.10, 1.30), 4 * (3 + 3 * 4) + 2),    # bottom-right corner: bottom
        ((1.30, 1.10), 4 * (3 + 3 * 4) + 1),    # bottom-right corner: right
        ((0.0, 0.5), 4 * (0 + 2 * 4) + 3),      # on the left wall counts as outside-left
        ((1.0, 0.5), 4 * (3 + 2 * 4) + 1),      # on the right wall
        ((0.5, 0.0), 4 * (2 + 0 * 4) + 0),      # on the top wall
        ((0.5, 1.0), 4 * (2 + 3 * 4) + 2),      # on the bottom wall
    ]
    pts = np.array([c[0] for c in cases])
    got = orc.calc_cell_idx(T.geom(), pts)
    assert list(got) == [c[1] for c in cases]


def test_2d_type_I_centre_diagonals(orc):
    # inside a rectangle, the two diagonals split it into top/right/bottom/left
    T = Tessellation([0, 0], [1, 1], (1, 1), "I")
    pts = np.array([[0.5, 0.1], [0.9, 0.5], [0.5, 0.9], [0.1, 0.5]])
    assert list(orc.calc_cell_idx(T.geom(), pts)) == [0, 1, 2, 3]


def test_nan_lookup(orc):
    # max(0, NaN) is 0, so a NaN coordinate is looked up at the origin: cell 0,
    # except for tetrahedra where the origin lies on -x-y+z = 0 (tetra 1).
    for dim, tess, want in ((1, 2, 0), (2, 1, 0), (2, 2, 0), (3, 1, 1), (3, 2, 0)):
        nCs = [4] * dim
        T = Tessellation([0] * dim, [1.0] * dim, nCs, "I" if tess == 1 else "II")
        p = np.full((1, dim), np.nan)
        assert orc.calc_cell_idx(T.geom(), p)[0] == want


def test_3d_type_I_z_clamp_uses_inc_x(orc):
    """calc_3Dto3D.cu:59 clamps z at nC2*inc_x - 1e-10.  With inc_x < inc_z a
    point above nC2*inc_x is looked up as if it sat at that height (a quirk
    of the reference kept for parity)."""
    T = Tessellation([0, 0, 0], [1.0, 1.0, 1.0], (3, 2, 2), "I")
    assert T.incs[0] < T.incs[2]
    top = 2 * T.incs[0] - 1e-10
    rng = np.random.default_rng(13)
    p = rng.uniform(0, 1, (500, 3))
    p[:, 2] = rng.uniform(top + 1e-6, 1.0, 500)
    q = p.copy()
    q[:, 2] = top
    assert np.array_equal(orc.calc_cell_idx(T.geom(), p), orc.calc_cell_idx(T.geom(), q))
