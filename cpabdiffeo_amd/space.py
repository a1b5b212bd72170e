"""CPA velocity-field spaces: constraints, basis, prior (host-side setup).

The reference builds the constraint matrix L (continuity across shared
facets, zero normal velocity on the boundary, optional volume preservation,
optional continuity of the outside extension) and takes B = null(L) by a
dense SVD (cpab/cpa2d/CpaSpace.py:86-219, cpab/cpaNd/utils/_null.py).  This
module builds the same constraint set but reaches the null space faster on
simplicial tessellations: continuous piecewise-affine fields on a conforming
triangulation / tetrahedralisation are exactly the piecewise-linear
interpolants of their vertex velocities, so the vertex "hat" fields span the
continuous space and only the remaining constraints need an SVD.

Any orthonormal basis of the null space is a valid B; theta is therefore
basis-dependent while the per-cell affine table (Avees) is not, which is why
parity fixtures store Avees.
"""
import numpy as np

from .tessellation import TESS_I, TESS_II, Tessellation


class PAT:
    """A piecewise-affine transformation: the per-cell affine matrices
    (cpab/cpaNd/_PAT.py:12-59).  ``As_sq_mats`` is the host (nC, n+1, n+1)
    form; the device table (nC, n*(n+1)) is uploaded by ``update``."""

    def __init__(self, pa_space, Avees=None):
        self.pa_space = pa_space
        self.nC = pa_space.nC
        n = pa_space.dim_domain
        self.As_sq_mats = np.zeros((self.nC, n + 1, n + 1))
        self._dev = None  # CpuGpuArray (nC, n*(n+1)), created on first device use
        self.Avees = np.zeros(pa_space.D)
        if Avees is not None:
            self.update(Avees)

    def update(self, Avees):
        """Copy Avees into the affine matrices (cpab/cpaNd/_PAT.py:48-59)."""
        sp = self.pa_space
        Avees = np.asarray(Avees, dtype=np.float64).reshape(-1)
        if sp.D != len(Avees):
            raise ValueError(sp.D, len(Avees))
        self.Avees[:] = Avees
        self.As_sq_mats[:, :-1, :] = Avees.reshape(self.nC, *sp.Ashape)
        if self._dev is not None:
            self._dev.cpu[:] = Avees.reshape(self.nC, -1)
            self._dev.cpu2gpu()

    def device_table(self):
        """(nC, n*(n+1)) float64 table on the GPU (the reference's As_vectorized)."""
        if self._dev is None:
            from .gpuarray import CpuGpuArray
            self._dev = CpuGpuArray(self.Avees.reshape(self.nC, -1))
        return self._dev

    def __repr__(self):
        return repr(self.As_sq_mats)


def _null(M, rtol=1e-10):
    """Orthonormal basis of the null space of M (columns)."""
    if M.shape[0] == 0:
        return np.eye(M.shape[1])
    u, s, vt = np.linalg.svd(M, full_matrices=True)
    tol = rtol * (s[0] if s.size else 1.0)
    rank = int((s > tol).sum())
    return vt[rank:].T.copy()


def _orth(M, rtol=1e-10):
    u, s, vt = np.linalg.svd(M, full_matrices=False)
    tol = rtol * (s[0] if s.size else 1.0)
    return u[:, s > tol].copy()


class CpaSpace:
    """A space of continuous piecewise-affine velocity fields on a box.

    Mirrors cpab/cpaNd/_CpaSpace.py and cpab/cpa{1,2,3}d/CpaSpace.py: the
    same constructor arguments (XMINS, XMAXS, nCs, zero_v_across_bdry,
    vol_preserve, tess, valid_outside) and the same public methods.  The
    integration methods run on the GPU through the HIP library.
    """

    my_dtype = np.float64

    def __init__(self, XMINS, XMAXS, nCs, zero_v_across_bdry=None, vol_preserve=False,
                 warp_around=None, zero_vals=(), tess="II", valid_outside=None,
                 only_local=False, cont_constraints_are_separable=False):
        self.tessellation = Tessellation(XMINS, XMAXS, nCs, tess)
        T = self.tessellation
        self.dim_domain = self.dim_range = dim = T.dim_domain
        self.nHomoCoo = dim + 1
        self.lengthAvee = dim * (dim + 1)
        self.Ashape = (dim, dim + 1)
        self.XMINS, self.XMAXS = T.XMINS, T.XMAXS
        self.nCs, self.nC, self.incs = T.nCs, T.nC, T.incs
        self.tess = "I" if T.type == TESS_I else "II"
        if zero_v_across_bdry is None:
            zero_v_across_bdry = [False] * dim
        if len(zero_v_across_bdry) != dim:
            raise ValueError(len(zero_v_across_bdry), dim)
        if warp_around is not None and any(warp_around):
            raise NotImplementedError("warp_around")
        self.zero_v_across_bdry = [bool(z) for z in zero_v_across_bdry]
        self.vol_preserve = bool(vol_preserve)
        if valid_outside is None:
            valid_outside = not any(self.zero_v_across_bdry)
        self.valid_outside = bool(valid_outside)
        # consistency rules of cpab/cpaNd/_CpaSpace.py:129-149
        if T.type == TESS_I and dim == 2:
            if any(self.zero_v_across_bdry) and self.valid_outside:
                raise ValueError("tess='I' with zero_v_across_bdry and valid_outside is inconsistent")
            if not all(self.zero_v_across_bdry) and not self.valid_outside:
                raise ValueError("tess='I' without zero_v_across_bdry needs valid_outside")
        elif T.type == TESS_I and dim == 3:
            if self.valid_outside:
                raise NotImplementedError("3D tess='I' with valid_outside")
            if not all(self.zero_v_across_bdry):
                raise ValueError("3D tess='I' needs zero_v_across_bdry")
        self.D = self.nC * self.lengthAvee
        self._build_basis(list(zero_vals))
        self.d = self.B.shape[1]
        if self.d == 0:
            raise ValueError("no degrees of freedom")
        # BasMats: (d, nC, dim, dim+1) (cpab/cpaNd/_CpaSpace.py:296-300)
        self.BasMats = self.B.T.reshape(self.d, self.nC, dim, dim + 1).copy()
        self.Avees = self.get_zeros_PA()
        self.As = self.Avees.reshape(self.nC, *self.Ashape)
        self.pat = PAT(self, self.get_zeros_PA())
        self._gpu = None  # device scratch, created on the first GPU call

    # ------------------------------------------------------------ basis
    def _vertices(self):
        verts = self.tessellation.cells_verts_homo_coo[:, :, :-1]
        flat = verts.reshape(-1, self.dim_domain)
        uniq, inv = np.unique(flat, axis=0, return_inverse=True)
        return uniq, inv.reshape(verts.shape[:2])

    def _vertex_basis(self):
        """Avees of the hat field of every (vertex, component): (D, nV*dim)."""
        dim, ne, nC = self.dim_domain, self.lengthAvee, self.nC
        X = self.tessellation.cells_verts_homo_coo  # (nC, dim+1, dim+1) for simplices/intervals
        uniq, cell_v = self._vertices()
        nV = len(uniq)
        B0 = np.zeros((self.D, nV * dim))
        for c in range(nC):
            Xinv = np.linalg.inv(X[c])  # A = V^T X^{-T}; hat of local vertex u is column u
            for u in range(dim + 1):
                g = cell_v[c, u]
                for r in range(dim):
                    B0[c * ne + r * (dim + 1): c * ne + (r + 1) * (dim + 1), g * dim + r] = Xinv[:, u]
        self._vert_tess = uniq
        self._cell_v = cell_v
        return B0

    def _constraint_rows(self, zero_vals):
        dim, ne = self.dim_domain, self.lengthAvee
        T = self.tessellation
        X = T.cells_verts_homo_coo
        rows = []

        def row_eval(c, r, v, coef=1.0, out=None):
            out = np.zeros(self.D) if out is None else out
            out[c * ne + r * (dim + 1): c * ne + (r + 1) * (dim + 1)] += coef * v
            return out

        simplicial = (T.type == TESS_I) or dim == 1
        if not simplicial:
            # continuity at the shared vertices of adjacent boxes
            keys = [set(map(tuple, X[c])) for c in range(self.nC)]
            for a in range(self.nC):
                for b in range(a + 1, self.nC):
                    shared = keys[a] & keys[b]
                    if len(shared) < 2 ** (dim - 1):
                        continue
                    for v in shared:
                        v = np.asarray(v)
                        for r in range(dim):
                            rows.append(row_eval(b, r, v, -1.0, row_eval(a, r, v)))
        if T.type == TESS_I and dim == 2 and self.valid_outside:
            # continuity of the outside extension: adjacent boundary triangles
            # agree on the line through their shared vertex, normal to the
            # wall (cpab/cpa2d/Tessellation.py:212-240, 300-322)
            nCx, nCy = T.nCs
            idx = {m: c for c, m in enumerate(T.cells_multiidx)}
            pairs = []
            for i in range(nCy - 1):
                pairs.append((idx[(0, i, 3)], idx[(0, i + 1, 3)], 0))
                pairs.append((idx[(nCx - 1, i, 1)], idx[(nCx - 1, i + 1, 1)], 0))
            for j in range(nCx - 1):
                pairs.append((idx[(j, 0, 0)], idx[(j + 1, 0, 0)], 1))
                pairs.append((idx[(j, nCy - 1, 2)], idx[(j + 1, nCy - 1, 2)], 1))
            for a, b, axis in pairs:
                shared = set(map(tuple, X[a])) & set(map(tuple, X[b]))
                (v,) = [np.asarray(s) for s in shared]
                aux = v.copy()
                aux[axis] -= 10
                for p in (v, aux):
                    for r in range(dim):
                        rows.append(row_eval(b, r, p, -1.0, row_eval(a, r, p)))
        for k, z in enumerate(self.zero_v_across_bdry):
            if not z:
                continue
            for c in range(self.nC):
                for v in X[c]:
                    if v[k] in (self.XMINS[k], self.XMAXS[k]):
                        rows.append(row_eval(c, k, v))
        if self.vol_preserve:
            for c in range(self.nC):
                row = np.zeros(self.D)
                for r in range(dim):
                    row[c * ne + r * (dim + 1) + r] = 1.0
                rows.append(row)
        for (r, col) in zero_vals:
            for c in range(self.nC):
                row = np.zeros(self.D)
                row[c * ne + r * (dim + 1) + col] = 1.0
                rows.append(row)
        return np.asarray(rows).reshape(-1, self.D)

    def _build_basis(self, zero_vals):
        T = self.tessellation
        simplicial = (T.type == TESS_I) or self.dim_domain == 1
        self.local_basis = self._vertex_basis() if simplicial else None
        B0 = self.local_basis if simplicial else np.eye(self.D)
        L = self._constraint_rows(zero_vals)
        self.constraintMat = L
        N = _null(L @ B0) if L.shape[0] else np.eye(B0.shape[1])
        self.B = _orth(B0 @ N)

    # ---------------------------------------------------- conversions
    def get_zeros_PA(self):
        return np.zeros(self.D)

    def get_zeros_theta(self):
        return np.zeros(self.d)

    def theta2Avees(self, theta, Avees=None):
        """Avees = B theta (cpab/cpaNd/_CpaSpace.py:366-376)."""
        if Avees is None:
            Avees = self.Avees
        np.dot(self.B, theta, out=Avees)
        return Avees

    def Avees2theta(self, Avees=None, theta=None):
        """theta = B^T Avees (a projection onto the space)."""
        if Avees is None:
            Avees = self.Avees
        if theta is None:
            return self.B.T.dot(Avees)
        np.dot(self.B.T, Avees, out=theta)
        return theta

    def Avees2As(self, Avees=None, As=None):
        if Avees is None:
            Avees = self.Avees
        out = np.asarray(Avees).reshape(self.nC, *self.Ashape)
        if As is None:
            return out.copy()
        As[:] = out
        return As

    def As2Avees(self, As, Avees=None):
        flat = np.asarray(As).reshape(-1)
        if Avees is None:
            return flat.copy()
        Avees[:] = flat
        return Avees

    def zeros_velTess(self):
        if self.local_basis is None:
            raise ValueError("vertex velocities need a simplicial tessellation")
        return np.zeros((len(self._vert_tess), self.dim_domain))

    def velTess2Avees(self, velTess, Avees=None):
        """Per-vertex velocities -> Avees (the reference's local version,
        cpab/cpaNd/utils/_get_stuff_for_the_local_version.py:43-106)."""
        if self.local_basis is None:
            raise ValueError("vertex velocities need a simplicial tessellation")
        out = self.local_basis @ np.asarray(velTess, np.float64).reshape(-1)
        if Avees is None:
            Avees = self.Avees
        Avees[:] = out
        return Avees

    def Avees2velTess(self, Avees=None, velTess=None):
        if self.local_basis is None:
            raise ValueError("vertex velocities need a simplicial tessellation")
        if Avees is None:
            Avees = self.Avees
        vals = np.zeros((len(self._vert_tess), self.dim_domain))
        As = np.asarray(Avees).reshape(self.nC, *self.Ashape)
        X = self.tessellation.cells_verts_homo_coo
        for c in range(self.nC):
            for u in range(self.dim_domain + 1):
                vals[self._cell_v[c, u]] = As[c] @ X[c, u]
        if velTess is None:
            return vals
        velTess[:] = vals
        return velTess

    @property
    def vert_tess(self):
        return self._vert_tess

    # ------------------------------------------------------------ prior
    def se_prior_cov(self, length_scale=0.1, variance=1.0):
        """Covariance of theta under a squared-exponential prior on Avees.

        Sigma_PA = K (x) I, K_ij = variance * exp(-|c_i - c_j|^2 / (2 l^2))
        over the cell centroids, and Sigma_CPA = B^T Sigma_PA B.  The length
        scale is in units of the largest domain extent."""
        cen = self.tessellation.cell_centroids() / float(np.max(self.XMAXS - self.XMINS))
        d2 = ((cen[:, None, :] - cen[None, :, :]) ** 2).sum(-1)
        K = variance * np.exp(-d2 / (2.0 * length_scale ** 2))
        Bc = self.B.reshape(self.nC, self.lengthAvee, self.d)
        cov = np.zeros((self.d, self.d))
        for e in range(self.lengthAvee):
            Be = Bc[:, e, :]
            cov += Be.T @ K @ Be
        return 0.5 * (cov + cov.T)

    def sample_theta(self, cov, rng, n=None):
        """theta ~ N(0, cov), drawn through an eigendecomposition."""
        w, U = np.linalg.eigh(cov)
        S = U * np.sqrt(np.clip(w, 0.0, None))
        if n is None:
            return S @ rng.standard_normal(self.d)
        return (S @ rng.standard_normal((self.d, n))).T.copy()

    # ------------------------------------------------------- transformation
    def update_pat(self, Avees=None):
        """cpab/cpaNd/_CpaSpace.py:438-441."""
        if Avees is None:
            Avees = self.Avees
        self.pat.update(Avees)

    def _dev(self):
        """Backend and per-space device scratch (cpab/cpaNd/_CpaSpace.py:190-204)."""
        if self._gpu is None:
            from .calcs import Calcs
            from .gpuarray import CpuGpuArray
            ne = self.lengthAvee
            self._gpu = {
                "calcs": Calcs(self.nC, np.float64, self.dim_domain, self.dim_range, self.tess),
                "signed_As": CpuGpuArray.zeros((self.nC, ne)),
                "Tlocals": CpuGpuArray.zeros((self.nC, ne)),
            }
        return self._gpu

    @property
    def _gpu_calcs(self):
        return self._dev()["calcs"]

    def _check_pts(self, pts, out=None, what="pts"):
        from .gpuarray import CpuGpuArray
        import torch
        for x in (pts, out):
            if x is None:
                continue
            if not isinstance(x, (CpuGpuArray, torch.Tensor)):
                raise TypeError(type(x))
            if x.ndim != 2 or x.shape[1] != self.dim_domain:
                raise ValueError(tuple(x.shape))
        if out is not None:
            if pts is out:
                raise ValueError("pts and out cannot be the same object")
            if tuple(out.shape) != tuple(pts.shape):
                raise ValueError(tuple(out.shape), tuple(pts.shape))

    def prepare_signed_As_and_Tlocals(self, mysign, dt, pat=None):
        """sign * As and expm(dt * sign * As) on the device
        (cpab/cpaNd/_CpaCalcs.py:227-313)."""
        from .calcs import gpu_prep_trels
        pat = self.pat if pat is None else pat
        g = self._dev()
        gpu_prep_trels(self.dim_domain, pat.device_table(), int(mysign), float(dt),
                       g["signed_As"], g["Tlocals"])
        return g["signed_As"], g["Tlocals"]

    def _calc_T(self, pat=None, pts=None, mysign=1, do_checks=True, out=None, timer=None,
                n_fallback=None, **params_flow_int):
        """cpab/cpaNd/_CpaSpace.py:632-667 -> _CpaCalcs.py:376-558."""
        if pts is None:
            raise ValueError
        if out is None:
            raise ValueError("You must pass the OUT argument")
        self._check_pts(pts, out)
        dt = params_flow_int["dt"]
        nT = params_flow_int["nTimeSteps"]
        nS = params_flow_int["nStepsODEsolver"]
        # prepare_signed_As_and_Tlocals and calc_T in one launch (the same tables and bits)
        pat = self.pat if pat is None else pat
        g = self._dev()
        self._gpu_calcs.calc_T_prep(pat.device_table(), int(mysign), pts, dt, nT, nS, out, self.dim_domain,
                                    self.dim_range, self.nCs, self.incs, g["signed_As"], g["Tlocals"],
                                    n_fallback=n_fallback)
        return out

    def calc_T_fwd(self, pat=None, pts=None, do_checks=True, out=None, **params_flow_int):
        return self._calc_T(pat=pat, pts=pts, mysign=1, do_checks=do_checks, out=out, **params_flow_int)

    def calc_T_inv(self, pat=None, pts=None, do_checks=True, out=None, **params_flow_int):
        return self._calc_T(pat=pat, pts=pts, mysign=-1, do_checks=do_checks, out=out, **params_flow_int)

    def calc_T_simple(self, pat=None, pts=None, mysign=1, do_checks=True, out=None, timer=None,
                      **params_flow_int):
        """cpab/cpaNd/_CpaSpace.py:590-624 -> _CpaCalcs.py:561-743."""
        if pts is None:
            raise ValueError
        if out is None:
            raise ValueError("You must pass the OUT argument")
        self._check_pts(pts, out)
        from .calcs import gpu_prep_trels
        pat = self.pat if pat is None else pat
        g = self._dev()
        # signed As only: the Trels this writes are not used by calc_T_simple
        gpu_prep_trels(self.dim_domain, pat.device_table(), int(mysign), 0.0, g["signed_As"], g["Tlocals"])
        self._gpu_calcs.calc_T_simple(None, None, g["signed_As"], pts, params_flow_int["dt"],
                                      params_flow_int["nTimeSteps"], params_flow_int["nStepsODEsolver"],
                                      out, self.dim_domain, self.dim_range, self.nCs, self.incs, timer=timer)
        return out

    def calc_grad_theta(self, pat=None, pts=None, mysign=1, do_checks=False, transformed=None,
                        grad_theta=None, grad_per_point=None, **params_flow_int):
        """cpab/cpaNd/_CpaSpace.py:669-686: calc_T_simple end points into
        `transformed` and dT/dtheta (nPts, dim, d) into `grad_per_point`.
        With mysign = -1 the field is -v but the basis tables are used as
        they are, as in the reference's kernel call (_CpaCalcs.py:770-792)."""
        if pts is None or transformed is None or grad_per_point is None:
            raise ValueError
        self._check_pts(pts, transformed)
        from .calcs import gpu_prep_trels
        from .gpuarray import CpuGpuArray
        pat = self.pat if pat is None else pat
        g = self._dev()
        gpu_prep_trels(self.dim_domain, pat.device_table(), int(mysign), 0.0, g["signed_As"], g["Tlocals"])
        if "BasMats" not in g:
            g["BasMats"] = CpuGpuArray(self.BasMats)
        self._gpu_calcs.calc_grad_theta(None, None, g["signed_As"], g["BasMats"], pts, params_flow_int["dt"],
                                        params_flow_int["nTimeSteps"], params_flow_int["nStepsODEsolver"],
                                        transformed, grad_per_point, self.dim_domain, self.dim_range,
                                        self.nCs, self.incs)
        return grad_per_point

    def calc_v(self, pat=None, pts=None, out=None, do_checks=True):
        """cpab/cpaNd/_CpaSpace.py:573-587 -> _CpaCalcs.py:95-165."""
        if pts is None:
            raise ValueError("Pts cannot be None")
        if out is None:
            raise ValueError("out Can't have None")
        self._check_pts(pts, out)
        pat = self.pat if pat is None else pat
        self._gpu_calcs.calc_velocities(None, None, pat.device_table(), pts, out, self.dim_domain,
                                        self.dim_range, self.nCs, self.incs)
        return out

    def calc_cell_idx(self, pts, cell_idx):
        """cpab/cpaNd/_CpaSpace.py:567-569 -> _CpaCalcs.py:59-88."""
        self._check_pts(pts)
        return self._gpu_calcs.calc_cell_idx(None, None, pts, cell_idx, self.dim_domain,
                                             self.dim_range, self.nCs, self.incs)

    def calc_trajectory(self, pat=None, pts=None, mysign=1, **params_flow_int):
        """cpab/cpaNd/_CpaSpace.py:688-703 -> _CpaCalcs.py:795-871.  Returns a
        CpuGpuArray (1 + nTimeSteps, nPts, dim) whose row 0 is pts."""
        if pts is None:
            raise ValueError("pts cannot be None")
        self._check_pts(pts)
        from .gpuarray import CpuGpuArray, dev_tensor
        import torch
        dt = params_flow_int["dt"]
        nT = int(params_flow_int["nTimeSteps"])
        nS = params_flow_int.get("nStepsODEsolver", 100)
        sA, T = self.prepare_signed_As_and_Tlocals(mysign, dt, pat)
        traj = self._gpu_calcs.calc_trajectory(None, None, T, sA, pts, dt, nT, nS, self.dim_domain,
                                               self.dim_range, self.nCs, self.incs)
        p0 = dev_tensor(pts, "float64")
        return CpuGpuArray.from_gpu(torch.cat([p0[None], traj], dim=0))

    def __repr__(self):
        return "CpaSpace(dim=%d, tess=%s, nCs=%s, nC=%d, d=%d)" % (
            self.dim_domain, self.tess, list(self.nCs), self.nC, self.d)
