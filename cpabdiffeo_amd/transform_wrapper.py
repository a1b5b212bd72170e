"""The public TransformWrapper surface (drop-in for cpab/cpaNd/_TransformWrapper.py
and cpab/cpa{1,2,3}d/TransformWrapper.py), backed by the HIP library.

Kept from the reference: constructor arguments, the pixel-unit domain
[0, nCols] x [0, nRows] (x [0, nSlices]), levels with base * 2**level cells
per axis, the integration parameters (normal / fine / coarse, selected by
``int_quality`` in {0, 1, -1}), method names and argument order, and the
exception types.  Arrays are ``CpuGpuArray`` (or ROCm torch tensors).

Not kept: plotting helpers, OpenCV remaps, the multiscale coarse-to-fine
prior (each level here has its own squared-exponential prior, scaled by
``scale_spatial`` and ``scale_value``).
"""
import copy

import numpy as np

from .gpuarray import CpuGpuArray, dev_tensor, torch_mod
from .space import CpaSpace


class Bunch(dict):
    def __getattr__(self, k):
        try:
            return self[k]
        except KeyError:
            raise AttributeError(k)

    def __setattr__(self, k, v):
        self[k] = v


def default_params_flow_int(dim):
    """cpab/cpa2d/calcs/_default_params_flow_int.py:11-14 (2D and 3D) and
    cpab/cpa1d/calcs/_default_params_flow_int.py:11-14 (1D)."""
    if dim == 1:
        b = Bunch(dt=0.01)
        b.nTimeSteps = int(1.0 / b.dt)
        b.nStepsODEsolver = 10
        return b
    return Bunch(nTimeSteps=10, dt=0.0001, nStepsODEsolver=10)


class ObsoleteError(Exception):
    pass


class TransformWrapperNd:
    """Dimension-independent part (cpab/cpaNd/_TransformWrapper.py:12-287)."""

    dim_domain = None

    def _init_common(self, XMAXS, vol_preserve, nLevels, base, scale_spatial, scale_value,
                     zero_v_across_bdry, tess, valid_outside, only_local):
        if not isinstance(vol_preserve, bool):
            raise TypeError(type(vol_preserve), "expected bool")
        if valid_outside is not None and not isinstance(valid_outside, bool) and self.dim_domain > 1:
            raise TypeError(type(valid_outside), "expected bool")
        if tess == "tri":
            tess = "I"
        if tess == "rect":
            tess = "II"
        self.nLevels = nLevels
        self.args = Bunch(vol_preserve=vol_preserve, nLevels=nLevels, base=base,
                          scale_spatial=scale_spatial, scale_value=scale_value,
                          zero_v_across_bdry=zero_v_across_bdry, tess=tess,
                          valid_outside=valid_outside, only_local=only_local)
        base = np.asarray(base)
        self.L_cpa_space = []
        for i in range(nLevels):
            nCs = base * 2 ** i  # cpab/cpa2d/Multiscale.py:71-79
            self.L_cpa_space.append(CpaSpace([0] * self.dim_domain, XMAXS, list(nCs),
                                             zero_v_across_bdry=zero_v_across_bdry,
                                             vol_preserve=vol_preserve, tess=tess,
                                             valid_outside=valid_outside))
        self.ms = self  # the reference reaches the spaces through tw.ms.L_cpa_space
        self.XMINS = np.zeros(self.dim_domain)
        self.XMAXS = np.asarray(XMAXS, dtype=np.float64)
        self._prior_cov = [None] * nLevels
        self.params_flow_int = default_params_flow_int(self.dim_domain)
        self.params_flow_int.nStepsODEsolver = 10
        self.params_flow_int_coarse = copy.deepcopy(self.params_flow_int)
        self.params_flow_int_coarse.nTimeSteps = int(self.params_flow_int_coarse.nTimeSteps // 10)
        self.params_flow_int_coarse.dt *= 10
        self.params_flow_int_fine = copy.deepcopy(self.params_flow_int)
        self.params_flow_int_fine.nTimeSteps *= 10
        self.params_flow_int_fine.dt /= 10
        self.timer = Bunch()

    # ------------------------------------------------------------ levels
    def try_to_determine_level(self):
        if self.nLevels == 1:
            return 0
        raise ValueError("Can't determine level: there are %d levels" % self.nLevels)

    @property
    def pats(self):
        return [s.pat for s in self.L_cpa_space]

    def get_zeros_theta(self, level):
        return self.L_cpa_space[level].get_zeros_theta()

    def get_zeros_theta_all_levels(self):
        return [s.get_zeros_theta() for s in self.L_cpa_space]

    def get_zeros_PA(self, level):
        return self.L_cpa_space[level].get_zeros_PA()

    def get_zeros_PA_all_levels(self):
        return [s.get_zeros_PA() for s in self.L_cpa_space]

    # ------------------------------------------------------------- prior
    def prior_cov(self, level):
        if self._prior_cov[level] is None:
            a = self.args
            self._prior_cov[level] = self.L_cpa_space[level].se_prior_cov(
                length_scale=a.scale_spatial, variance=float(a.scale_value))
        return self._prior_cov[level]

    def sample_gaussian(self, level, Avees, theta, mu, rng=None):
        """theta ~ N(mu, Sigma_CPA), Avees = B theta (modifies both in place)."""
        rng = np.random.default_rng() if rng is None else rng
        sp = self.L_cpa_space[level]
        theta[:] = mu + sp.sample_theta(self.prior_cov(level), rng)
        sp.theta2Avees(theta, Avees)

    def sample_from_the_ms_prior_coarse2fine_all_levels(self, rng=None):
        """Independent per-level prior draws: returns (ms_Avees, ms_theta)."""
        ms_Avees = self.get_zeros_PA_all_levels()
        ms_theta = self.get_zeros_theta_all_levels()
        for lev in range(self.nLevels):
            self.sample_gaussian(lev, ms_Avees[lev], ms_theta[lev], 0.0, rng)
        return ms_Avees, ms_theta

    # -------------------------------------------------------------- pats
    def update_pat_from_Avees(self, Avees=None, level=None):
        """cpab/cpaNd/_TransformWrapper.py:109-112."""
        if level is None:
            level = self.try_to_determine_level()
        sp = self.L_cpa_space[level]
        if Avees is None:
            Avees = sp.Avees
        sp.update_pat(Avees)

    def update_ms_pats(self, ms_Avees=None):
        if ms_Avees is None:
            raise NotImplementedError
        for sp, Av in zip(self.L_cpa_space, ms_Avees):
            sp.update_pat(Av)

    def update_pat_from_velTess(self, velTess, level=None):
        if level is None:
            raise ValueError("I am not going to try guess what level you want")
        sp = self.L_cpa_space[level]
        sp.velTess2Avees(velTess)
        sp.update_pat()

    # ------------------------------------------------------------- calcs
    def calc_cell_idx(self, pts_src, cell_idx, level, permute_for_disp=False):
        """cpab/cpaNd/_TransformWrapper.py:125-136."""
        sp = self.L_cpa_space[level]
        sp.calc_cell_idx(pts_src, cell_idx)
        if permute_for_disp:
            raise NotImplementedError("display permutation")

    def calc_v(self, level, pts=None, v=None):
        """cpab/cpaNd/_TransformWrapper.py:141-158."""
        if pts is None and v is None:
            pts, v = self.pts_src_dense, self.v_dense
        elif pts is not None and v is not None:
            if tuple(pts.shape) != tuple(v.shape):
                raise ValueError(pts.shape, v.shape)
        else:
            raise ValueError("pts and v must be either both None or both not None")
        self.L_cpa_space[level].calc_v(pts=pts, out=v)

    def calc_T(self, *args, **kwargs):
        raise ObsoleteError("Use calc_T_inv or calc_T_fwd instead")

    def _params(self, int_quality):
        if int_quality not in (-1, 0, 1):
            raise ValueError
        return [self.params_flow_int, self.params_flow_int_fine, self.params_flow_int_coarse][int_quality]

    def _calc_T(self, pts_src, pts_transformed, mysign, level, timeit=False, int_quality=0):
        """cpab/cpaNd/_TransformWrapper.py:207-253."""
        p = self._params(int_quality)
        self.L_cpa_space[level]._calc_T(pts=pts_src, mysign=mysign, out=pts_transformed, **p)

    def _calc_T_simple(self, pts_src, pts_transformed, mysign, level, timeit=False, int_quality=0):
        p = self._params(int_quality)
        self.L_cpa_space[level].calc_T_simple(pts=pts_src, mysign=mysign, out=pts_transformed, **p)

    def calc_T_fwd(self, pts_src, pts_fwd, level, timeit=False, int_quality=0):
        self._calc_T(pts_src, pts_fwd, 1, level, timeit, int_quality)

    def calc_T_inv(self, pts_src, pts_inv, level, timeit=False, int_quality=0):
        self._calc_T(pts_src, pts_inv, -1, level, timeit, int_quality)

    def calc_T_fwd_simple(self, pts_src, pts_fwd, level, timeit=False, int_quality=0):
        self._calc_T_simple(pts_src, pts_fwd, 1, level, timeit, int_quality)

    def calc_T_inv_simple(self, pts_src, pts_fwd, level, timeit=False, int_quality=0):
        self._calc_T_simple(pts_src, pts_fwd, -1, level, timeit, int_quality)

    def calc_trajectory(self, pts=None, mysign=1, level=None):
        """cpab/cpa2d/TransformWrapper.py:213-226: (1 + nTimeSteps, nPts, dim)."""
        if level is None:
            level = self.try_to_determine_level()
        if pts is None:
            raise ValueError("pts can't be None")
        if not isinstance(pts, CpuGpuArray) and not hasattr(pts, "is_cuda"):
            raise TypeError(type(pts))
        return self.L_cpa_space[level].calc_trajectory(pts=pts, mysign=mysign, **self.params_flow_int)

    def remap(self, *args, **kwargs):
        raise ObsoleteError("Use remap_inv or remap_fwd instead")

    def __repr__(self):
        return "tw:\n\t" + "\n\t".join(repr(s) for s in self.L_cpa_space)


def _img(x, ndim):
    t = dev_tensor(x, "float32")
    if t.ndim != ndim:
        raise ValueError(tuple(t.shape))
    return t


class TransformWrapper(TransformWrapperNd):
    """2D wrapper (cpab/cpa2d/TransformWrapper.py:22-135)."""

    dim_domain = 2

    def __init__(self, nRows, nCols, vol_preserve=False, nLevels=1, base=(2, 2),
                 scale_spatial=1.0 * .1, scale_value=100, zero_v_across_bdry=(False, False),
                 tess=None, valid_outside=True, only_local=False,
                 cont_constraints_are_separable=False):
        if tess in ("tri",):
            tess = "I"
        if tess in ("rect",):
            tess = "II"
        if tess not in ("I", "II"):
            raise ValueError(tess, "tess must be in ['I','II']")
        if only_local and tess != "I":
            raise NotImplementedError
        self.nRows, self.nCols = int(nRows), int(nCols)
        self._init_common([nCols, nRows], vol_preserve, nLevels, base, scale_spatial, scale_value,
                          list(zero_v_across_bdry), tess, valid_outside, only_local)
        self._dense = None

    def _dense_arrays(self):
        if self._dense is None:
            pts = self.L_cpa_space[0].tessellation.dense_grid_img()
            self._dense = (CpuGpuArray(pts), CpuGpuArray.zeros(pts.shape),
                           CpuGpuArray.zeros(pts.shape))
        return self._dense

    @property
    def pts_src_dense(self):
        return self._dense_arrays()[0]

    @property
    def v_dense(self):
        return self._dense_arrays()[1]

    @property
    def transformed_dense(self):
        return self._dense_arrays()[2]

    def remap_fwd(self, pts_inv, img, img_wrapped_fwd, timeit=False):
        """img_wrapped_fwd(x, y) = img(T^{-1}(x, y)) (cpab/cpa2d/TransformWrapper.py:261-271)."""
        self._remap(pts_inv, img, img_wrapped_fwd)

    def remap_inv(self, pts_fwd, img, img_wrapped_inv, timeit=False):
        """img_wrapped_inv(x, y) = img(T(x, y)) (cpab/cpa2d/TransformWrapper.py:273-283)."""
        self._remap(pts_fwd, img, img_wrapped_inv)

    def _remap(self, pts_inv_or_fwd, img, img_wrapped):
        from .calcs import context_for
        if not isinstance(pts_inv_or_fwd, CpuGpuArray) and not hasattr(pts_inv_or_fwd, "is_cuda"):
            raise ObsoleteError
        src = _img(img, 2)
        out = _img(img_wrapped, 2)
        H, W = src.shape
        if tuple(out.shape) != (H, W):
            raise ValueError(tuple(out.shape), (H, W))
        pts = dev_tensor(pts_inv_or_fwd, "float64", (H * W, 2))
        ctx = context_for(src)
        ctx.call("cpab_remap2d", pts.data_ptr(), src.data_ptr(), out.data_ptr(), H, W)


class TransformWrapper1D(TransformWrapperNd):
    """1D wrapper (cpab/cpa1d/TransformWrapper.py:21-103): domain [0, nCols]."""

    dim_domain = 1

    def __init__(self, nCols, vol_preserve=False, nLevels=1, base=(5,), scale_spatial=1.0 * 10,
                 scale_value=2, zero_v_across_bdry=(True,), nPtsDense=None, only_local=False):
        self.nCols = nCols
        self._init_common([nCols], vol_preserve, nLevels, base, scale_spatial, scale_value,
                          list(zero_v_across_bdry), "II", None, only_local)
        self.nPtsDense = nPtsDense


class TransformWrapper3D(TransformWrapperNd):
    """3D wrapper (cpab/cpa3d/TransformWrapper.py:22-297): domain
    [0, nCols] x [0, nRows] x [0, nSlices]."""

    dim_domain = 3

    def __init__(self, nRows, nCols, nSlices, vol_preserve=False, nLevels=1, base=(2, 2, 2),
                 scale_spatial=1.0 * .1, scale_value=100, zero_v_across_bdry=(True, True, True),
                 tess=None, valid_outside=False, only_local=False, cont_constraints_are_separable=True):
        if tess not in ("II", "I"):
            raise ValueError(tess, "tess must be in ['II','I']")
        self.nRows, self.nCols, self.nSlices = int(nRows), int(nCols), int(nSlices)
        self._init_common([nCols, nRows, nSlices], vol_preserve, nLevels, base, scale_spatial,
                          scale_value, list(zero_v_across_bdry), tess, valid_outside, only_local)

    def remap_fwd(self, pts_inv, img, img_wrapped_fwd, timeit=False):
        self._remap(pts_inv, img, img_wrapped_fwd)

    def remap_inv(self, pts_fwd, img, img_wrapped_inv, timeit=False):
        self._remap(pts_fwd, img, img_wrapped_inv)

    def _remap(self, pts, vol, out):
        """Trilinear remap of a (nSlices, nRows, nCols) float volume
        (cpab/cpa3d/TransformWrapper.py:254-290)."""
        from .calcs import context_for
        src = _img(vol, 3)
        dst = _img(out, 3)
        D, H, W = src.shape
        if tuple(dst.shape) != (D, H, W):
            raise ValueError(tuple(dst.shape), (D, H, W))
        p = dev_tensor(pts, "float64", (D * H * W, 3))
        ctx = context_for(src)
        ctx.call("cpab_remap3d", p.data_ptr(), src.data_ptr(), dst.data_ptr(), D, H, W)
