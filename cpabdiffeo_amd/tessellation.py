"""Tessellations of a box domain into CPA cells (host-side geometry).

Cell order, vertex order and the increments passed to the kernels follow the
reference, because the kernels' cell lookup returns indices in that order:

* 1D intervals, left to right (cpab/cpa1d/Tessellation.py).
* 2D type II rectangles, row-major with x fastest; type I splits each
  rectangle into 4 triangles around its centre in the order top, right,
  bottom, left (cpab/cpa2d/Tessellation.py:83-126).
* 3D type II boxes, x fastest then y then z; type I splits each box into 5
  tetrahedra (a central one and four corners) whose diagonal orientation
  alternates with the parity of (i, j, k) (cpab/cpa3d/Tessellation.py:85-174).
* ``incs[i]`` is the first difference of the sorted distinct cell-centre
  coordinates along axis i, or twice the only centre when there is one cell
  along that axis (cpab/cpaNd/_CpaSpace.py:278-294).  For 50 cells on [0, 1]
  this gives 0.019999999999999997, not 0.02, and the kernels see exactly that.
"""
import numpy as np

TESS_I = 1   # triangles / tetrahedra (the reference's TESS_TYPE 1)
TESS_II = 2  # intervals / rectangles / boxes (TESS_TYPE 2)


def tess_code(tess):
    """Map 'I'/'II' to the kernels' TESS_TYPE (cpab/gpu/Calcs.py:161)."""
    if tess in ("I", "tri", TESS_I):
        return TESS_I
    if tess in ("II", "rect", TESS_II, None):
        return TESS_II
    raise ValueError(tess, "tess must be in ['I','II']")


class Tessellation:
    """Cells of ``[XMINS, XMAXS]`` split ``nCs`` times per axis.

    Attributes mirror the reference: ``cells_verts_homo_coo`` (nC, nV, dim+1),
    ``cells_multiidx``, ``box_centers`` (one row per rectangle/box, homogeneous),
    ``nC``, ``nCs``, ``incs``.
    """

    def __init__(self, XMINS, XMAXS, nCs, tess="II"):
        self.XMINS = np.asarray(XMINS, dtype=np.float64)
        self.XMAXS = np.asarray(XMAXS, dtype=np.float64)
        self.dim_domain = dim = len(self.XMINS)
        if dim not in (1, 2, 3):
            raise NotImplementedError(dim)
        if len(self.XMAXS) != dim or len(nCs) != dim:
            raise ValueError(XMINS, XMAXS, nCs)
        if (self.XMINS >= self.XMAXS).any():
            raise ValueError(XMINS, XMAXS)
        if self.XMINS.any():
            # The kernels' lookup assumes the domain starts at the origin
            # (cpab/cpa2d/calcs/_CpaCalcs.py:38-39, cpab/cpaNd/_Multiscale.py:74-75).
            raise NotImplementedError("XMINS must be 0")
        self.type = tess_code(tess)
        if self.type == TESS_I and dim == 1:
            raise ValueError("tess 'I' needs dim 2 or 3")
        self.nCs = np.asarray([int(n) for n in nCs])
        if (self.nCs < 1).any():
            raise ValueError(nCs)
        nboxes = int(np.prod(self.nCs))
        self.nC = nboxes * (1 if self.type == TESS_II else (4 if dim == 2 else 5))
        self.V = [np.linspace(self.XMINS[i], self.XMAXS[i], self.nCs[i] + 1) for i in range(dim)]
        self.cells_multiidx, self.cells_verts_homo_coo, self.box_centers = self._create_cells()
        self.incs = self._incs()
        self._xmins = self.cells_verts_homo_coo[:, :, :-1].min(axis=1)
        self._xmaxs = self.cells_verts_homo_coo[:, :, :-1].max(axis=1)

    # ------------------------------------------------------------- cells
    def _create_cells(self):
        dim, t = self.dim_domain, self.type
        cells, verts = [], []
        if dim == 1:
            Vx = self.V[0]
            for j in range(self.nCs[0]):
                cells.append((j,))
                verts.append(((Vx[j], 1.0), (Vx[j + 1], 1.0)))
        elif dim == 2:
            Vx, Vy = self.V
            for i in range(self.nCs[1]):
                for j in range(self.nCs[0]):
                    ul = (Vx[j], Vy[i], 1.0)
                    ur = (Vx[j + 1], Vy[i], 1.0)
                    ll = (Vx[j], Vy[i + 1], 1.0)
                    lr = (Vx[j + 1], Vy[i + 1], 1.0)
                    center = ((Vx[j] + Vx[j + 1]) / 2, (Vy[i] + Vy[i + 1]) / 2, 1.0)
                    if t == TESS_II:
                        cells.append((j, i))
                        verts.append((ul, ur, lr, ll))
                    else:
                        for k, (a, b) in enumerate(((ul, ur), (ur, lr), (lr, ll), (ll, ul))):
                            cells.append((j, i, k))
                            verts.append((center, a, b))
        else:
            Vx, Vy, Vz = self.V
            for k in range(self.nCs[2]):
                for i in range(self.nCs[1]):
                    for j in range(self.nCs[0]):
                        ul0, ur0 = (Vx[j], Vy[i], Vz[k], 1.0), (Vx[j + 1], Vy[i], Vz[k], 1.0)
                        ll0, lr0 = (Vx[j], Vy[i + 1], Vz[k], 1.0), (Vx[j + 1], Vy[i + 1], Vz[k], 1.0)
                        ul1, ur1 = (Vx[j], Vy[i], Vz[k + 1], 1.0), (Vx[j + 1], Vy[i], Vz[k + 1], 1.0)
                        ll1, lr1 = (Vx[j], Vy[i + 1], Vz[k + 1], 1.0), (Vx[j + 1], Vy[i + 1], Vz[k + 1], 1.0)
                        if t == TESS_II:
                            cells.append((j, i, k))
                            verts.append((ul0, ur0, lr0, ll0, ul1, ur1, lr1, ll1))
                            continue
                        # orientation flip on alternating boxes (cpab/cpa3d/Tessellation.py:150-160)
                        if k % 2 == 0:
                            flip = (i % 2 == 0 and j % 2 == 1) or (i % 2 == 1 and j % 2 == 0)
                        else:
                            flip = (i % 2 == 0 and j % 2 == 0) or (i % 2 == 1 and j % 2 == 1)
                        if flip:
                            ul0, ur0, lr0, ll0 = ur0, lr0, ll0, ul0
                            ul1, ur1, lr1, ll1 = ur1, lr1, ll1, ul1
                        # central tetrahedron first, then the four corners
                        for l, tet in enumerate(((ll1, ur1, ul0, lr0), (ul1, ur1, ll1, ul0),
                                                 (lr1, ur1, ll1, lr0), (ll0, ul0, lr0, ll1),
                                                 (ur0, ul0, lr0, ur1))):
                            cells.append((j, i, k, l))
                            verts.append(tet)
        verts = np.asarray(verts, dtype=np.float64)
        if len(verts) != self.nC:
            raise ValueError(len(verts), self.nC)
        # The box centres are computed as the reference computes them (numpy
        # means over the same vertex lists, so the same rounding), because
        # the increments the kernels divide by are derived from them.
        if t == TESS_I and dim == 2:
            centers = verts[::4][:, 0].copy()        # cpab/cpa2d/Tessellation.py:48
        elif t == TESS_I and dim == 3:
            centers = verts[::5].mean(axis=1)        # cpab/cpa3d/Tessellation.py:40
        else:
            centers = verts.mean(axis=1)             # cpa1d/Tessellation.py:30, cpa2d:51, cpa3d:42
        return cells, verts, centers

    def _incs(self):
        incs = []
        for i in range(self.dim_domain):
            col = self.box_centers[:, i]
            if self.nCs[i] > 1:
                incs.append(np.diff(np.unique(col))[0])
            else:
                incs.append(2 * np.unique(col)[0])
        incs = np.asarray(incs, dtype=np.float64)
        if (incs <= 0).any():
            raise ValueError(incs)
        nominal = (self.XMAXS - self.XMINS) / self.nCs
        if (np.abs(incs - nominal) > 1e-9 * nominal).any():
            # Centres that differ by an ulp make np.unique split them and the
            # first difference collapse; refuse instead of integrating garbage.
            raise ValueError("cell increments %s deviate from the cell size %s" % (incs, nominal))
        return incs

    # ----------------------------------------------------------- helpers
    def cell_centroids(self):
        """Centroid of every cell (mean of its vertices), shape (nC, dim)."""
        return self.cells_verts_homo_coo[:, :, :-1].mean(axis=1)

    def kernel_params(self):
        """(nC0, nC1, nC2, inc_x, inc_y, inc_z) as Calcs.parse_nCs_and_incs
        (cpab/gpu/Calcs.py:215-233) passes them to the kernels."""
        d = self.dim_domain
        nCs = [int(x) for x in self.nCs] + [0] * (3 - d)
        incs = [float(x) for x in self.incs] + [1.0] * (3 - d)
        return nCs[0], nCs[1], nCs[2], incs[0], incs[1], incs[2]

    def geom(self):
        """Geometry dict understood by the kernels' wrappers."""
        return {"dim": self.dim_domain, "tess": self.type,
                "nCs": [int(x) for x in self.nCs], "incs": [float(x) for x in self.incs]}

    def dense_grid_img(self):
        """Integer-pixel grid over the domain, row-major with x fastest:
        the reference's ``x_dense_img`` (cpab/cpa2d/calcs/_CpaCalcs.py:46-62)."""
        n = [int(round(x)) for x in self.XMAXS]
        if self.dim_domain == 1:
            return np.arange(n[0], dtype=np.float64)[:, None].copy()
        if self.dim_domain == 2:
            yy, xx = np.mgrid[0:n[1], 0:n[0]]
            return np.stack([xx.ravel(), yy.ravel()], axis=1).astype(np.float64)
        zz, yy, xx = np.mgrid[0:n[2], 0:n[1], 0:n[0]]
        return np.stack([xx.ravel(), yy.ravel(), zz.ravel()], axis=1).astype(np.float64)
