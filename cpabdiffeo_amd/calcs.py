"""The GPU backend object of a CPA space, mirroring the reference's ``Calcs``
(cpab/gpu/Calcs.py:44-1024) and ``gpu_expm``
(cpab/gpu/expm/expm_affine_by_series_2D.py:246-258), with the same method
names and argument meanings, launching the precompiled gfx950 kernels of
libcpab_hip.so through its C ABI.

Differences from the reference, by design:
 * no JIT and no shared-memory variants (``sharedmemory`` is accepted and
   ignored: the library stages tables in LDS whenever they fit);
 * float64 only (the reference's float32 rewrite of the kernel text is
   documented there as numerically unusable, Calcs.py:154-157);
 * shapes, dtypes and aliasing are always checked (the reference forces
   ``do_checks`` off on its hot path, Calcs.py:462).
"""
import numpy as np

from ._lib import Context
from .gpuarray import CpuGpuArray, dev_tensor, require_gpu, torch_mod
from .tessellation import tess_code


def _stream_bind(ctx, tensor):
    torch = torch_mod()
    ctx.bind_stream(torch.cuda.current_stream(tensor.device).cuda_stream)


def context_for(tensor):
    ctx = Context.get(tensor.device.index if tensor.device.index is not None else 0)
    _stream_bind(ctx, tensor)
    return ctx


class Calcs:
    """Backend of one CPA space (cpab/gpu/Calcs.py:44-63)."""

    def __init__(self, nCells, my_dtype, dim_domain, dim_range, tess, sharedmemory=2, device=None):
        if my_dtype not in (np.float32, np.float64):
            raise ValueError(my_dtype)
        if my_dtype != np.float64:
            raise NotImplementedError("the integration kernels are float64")
        if dim_range not in (1, dim_domain):
            raise NotImplementedError(dim_range, dim_domain)
        if dim_range != dim_domain:
            raise NotImplementedError("scalar-range CPA fields are outside this library")
        if tess not in ("I", "II"):
            raise NotImplementedError(tess)
        if tess == "I" and dim_domain not in (2, 3):
            raise NotImplementedError
        if sharedmemory not in (0, 1, 2):
            raise ValueError(sharedmemory)
        self.nCells = int(nCells)
        self.my_dtype = my_dtype
        self.dim_domain = dim_domain
        self.dim_range = dim_range
        self.tess = tess
        self.tess_type = tess_code(tess)
        self.sharedmemory = sharedmemory
        self.device = device

    @staticmethod
    def parse_nCs_and_incs(dim_domain, incs, nCs):
        """cpab/gpu/Calcs.py:215-233."""
        if dim_domain == 1:
            return int(nCs[0]), 0, 0, float(incs[0]), 1.0, 1.0
        if dim_domain == 2:
            return int(nCs[0]), int(nCs[1]), 0, float(incs[0]), float(incs[1]), 1.0
        if dim_domain == 3:
            return int(nCs[0]), int(nCs[1]), int(nCs[2]), float(incs[0]), float(incs[1]), float(incs[2])
        raise NotImplementedError(dim_domain)

    def _tables(self, dim, *tables):
        ne = dim * (dim + 1)
        return [dev_tensor(t, "float64", (self.nCells, ne)) for t in tables]

    def _pts(self, pts, dim):
        t = dev_tensor(pts, "float64")
        if t.ndim != 2 or t.shape[1] != dim:
            raise ValueError("Expected pts.ndim==2 with %d columns but pts.shape=%s" % (dim, tuple(t.shape)))
        return t

    # ------------------------------------------------------------ calc_T
    def calc_T(self, xmins, xmaxs, Tlocals_vectorized, signedAs_vectorized, pts0, dt, nTimeSteps,
               nStepsOdeSolver, pts_at_T, dim_domain, dim_range, nCs, incs, timer=None,
               n_fallback=None):
        """cpab/gpu/Calcs.py:454-561.  xmins/xmaxs are legacy and ignored, as
        by the reference kernels."""
        if pts_at_T is None:
            raise ValueError
        dim = dim_domain
        T, A = self._tables(dim, Tlocals_vectorized, signedAs_vectorized)
        p0 = self._pts(pts0, dim)
        out = self._pts(pts_at_T, dim)
        if tuple(out.shape) != tuple(p0.shape):
            raise ValueError(tuple(out.shape), tuple(p0.shape))
        nfb = 0
        if n_fallback is not None:
            nfb = dev_tensor(n_fallback, "int32", (p0.shape[0],)).data_ptr()
        ctx = context_for(p0)
        ctx.call("cpab_calc_T", dim, self.tess_type, self.nCells, p0.data_ptr(), out.data_ptr(),
                 T.data_ptr(), A.data_ptr(), float(dt), int(nTimeSteps), int(nStepsOdeSolver),
                 int(p0.shape[0]), *self.parse_nCs_and_incs(dim, incs, nCs), nfb)
        return pts_at_T

    def calc_T_prep(self, As_vectorized, mysign, pts0, dt, nTimeSteps, nStepsOdeSolver, pts_at_T, dim_domain,
                    dim_range, nCs, incs, signedAs_out, Tlocals_out, n_fallback=None):
        """prepare_signedSqAs_and_Tlocals_for_gpu + calc_T (cpab/cpaNd/_CpaCalcs.py:227-313, 376-558)
        in one launch: each block forms the signed table and the Trels from the raw As itself;
        the tables are also written to signedAs_out / Tlocals_out.  The same bits as the two
        calls."""
        if pts_at_T is None:
            raise ValueError
        if int(mysign) not in (1, -1):
            raise ValueError(mysign)
        dim = dim_domain
        A, sA, T = self._tables(dim, As_vectorized, signedAs_out, Tlocals_out)
        p0 = self._pts(pts0, dim)
        out = self._pts(pts_at_T, dim)
        if tuple(out.shape) != tuple(p0.shape):
            raise ValueError(tuple(out.shape), tuple(p0.shape))
        nfb = 0
        if n_fallback is not None:
            nfb = dev_tensor(n_fallback, "int32", (p0.shape[0],)).data_ptr()
        ctx = context_for(p0)
        ctx.call("cpab_calc_T_prep", dim, self.tess_type, self.nCells, p0.data_ptr(), out.data_ptr(), A.data_ptr(),
                 int(mysign), float(dt), int(nTimeSteps), int(nStepsOdeSolver), int(p0.shape[0]),
                 *self.parse_nCs_and_incs(dim, incs, nCs), sA.data_ptr(), T.data_ptr(), nfb)
        return pts_at_T

    def calc_T_simple(self, xmins, xmaxs, As_vectorized, pts0, dt, nTimeSteps, nStepsOdeSolver,
                      pts_at_T, dim_domain, dim_range, nCs, incs, timer=None):
        """cpab/gpu/Calcs.py:339-450."""
        if dim_domain != dim_range:
            raise ValueError
        if pts_at_T is None:
            raise ValueError
        if len(incs) != dim_domain:
            raise ValueError(len(incs), dim_domain)
        dim = dim_domain
        (A,) = self._tables(dim, As_vectorized)
        p0 = self._pts(pts0, dim)
        out = self._pts(pts_at_T, dim)
        if tuple(out.shape) != tuple(p0.shape):
            raise ValueError(tuple(out.shape), tuple(p0.shape))
        ctx = context_for(p0)
        ctx.call("cpab_calc_T_simple", dim, self.tess_type, self.nCells, p0.data_ptr(),
                 out.data_ptr(), A.data_ptr(), float(dt), int(nTimeSteps), int(nStepsOdeSolver),
                 int(p0.shape[0]), *self.parse_nCs_and_incs(dim, incs, nCs))
        return pts_at_T

    def calc_grad_theta(self, xmins, xmaxs, As_vectorized, BasMats, pts0, dt, nTimeSteps, nStepsOdeSolver,
                        pts_at_T, grad_per_point, dim_domain, dim_range, nCs, incs):
        """cpab/gpu/Calcs.py:237-335 (kernel calc_2Dto2D.cu:448-491): end points of
        calc_T_simple into pts_at_T and dT/dtheta into grad_per_point (nPts, dim, d)."""
        if dim_domain != dim_range:
            raise ValueError
        if pts_at_T is None:
            raise ValueError
        if len(incs) != dim_domain:
            raise ValueError(len(incs), dim_domain)
        dim = dim_domain
        (A,) = self._tables(dim, As_vectorized)
        Bs = dev_tensor(BasMats, "float64")
        d = Bs.shape[0]
        if tuple(Bs.shape) != (d, self.nCells, dim, dim + 1):
            raise ValueError(tuple(Bs.shape))
        p0 = self._pts(pts0, dim)
        out = self._pts(pts_at_T, dim)
        grad = dev_tensor(grad_per_point, "float64", (p0.shape[0], dim, d))
        ctx = context_for(p0)
        ctx.call("cpab_calc_grad_theta", dim, self.tess_type, self.nCells, p0.data_ptr(), out.data_ptr(),
                 A.data_ptr(), Bs.data_ptr(), grad.data_ptr(), int(d), float(dt), int(nTimeSteps),
                 int(nStepsOdeSolver), int(p0.shape[0]), *self.parse_nCs_and_incs(dim, incs, nCs))
        return grad_per_point

    def calc_trajectory(self, xmins, xmaxs, Tlocals_vectorized, As_vectorized, pts0, dt, nTimeSteps,
                        nStepsOdeSolver, dim_domain, dim_range, nCs, incs):
        """cpab/gpu/Calcs.py:735-814.  Returns a device tensor (nTimeSteps, nPts, dim):
        row t is the position after t+1 steps."""
        dim = dim_domain
        T, A = self._tables(dim, Tlocals_vectorized, As_vectorized)
        p0 = self._pts(pts0, dim)
        torch = torch_mod()
        traj = torch.empty((int(nTimeSteps), p0.shape[0], dim), dtype=torch.float64, device=p0.device)
        ctx = context_for(p0)
        ctx.call("cpab_calc_trajectory", dim, self.tess_type, self.nCells, p0.data_ptr(),
                 traj.data_ptr(), T.data_ptr(), A.data_ptr(), float(dt), int(nTimeSteps),
                 int(nStepsOdeSolver), int(p0.shape[0]), *self.parse_nCs_and_incs(dim, incs, nCs))
        return traj

    def calc_velocities(self, xmins, xmaxs, As_vectorized, pts, out, dim_domain, dim_range, nCs, incs):
        """cpab/gpu/Calcs.py:816-924."""
        dim = dim_domain
        (A,) = self._tables(dim, As_vectorized)
        p = self._pts(pts, dim)
        v = self._pts(out, dim)
        if tuple(v.shape) != tuple(p.shape):
            raise ValueError(tuple(v.shape), tuple(p.shape))
        ctx = context_for(p)
        ctx.call("cpab_calc_v", dim, self.tess_type, self.nCells, p.data_ptr(), v.data_ptr(),
                 A.data_ptr(), int(p.shape[0]), *self.parse_nCs_and_incs(dim, incs, nCs))
        return out

    def calc_cell_idx(self, xmins, xmaxs, pts, cell_idx, dim_domain, dim_range, nCs, incs):
        """cpab/gpu/Calcs.py:968-1010."""
        dim = dim_domain
        p = self._pts(pts, dim)
        c = dev_tensor(cell_idx, "int32", (p.shape[0],))
        ctx = context_for(p)
        ctx.call("cpab_calc_cell_idx", dim, self.tess_type, p.data_ptr(), c.data_ptr(),
                 int(p.shape[0]), *self.parse_nCs_and_incs(dim, incs, nCs))
        return cell_idx


def gpu_prep_trels(dim, As, sign, dt, signedAs_out, Trels_out):
    """Fused sign * As, dt scaling and expm (cpab/cpaNd/_CpaCalcs.py:227-313)."""
    A = dev_tensor(As, "float64")
    nC = A.shape[0]
    ne = dim * (dim + 1)
    if A.ndim != 2 or A.shape[1] != ne:
        raise ValueError(tuple(A.shape), ne)
    T = dev_tensor(Trels_out, "float64", (nC, ne))
    sA = dev_tensor(signedAs_out, "float64", (nC, ne)) if signedAs_out is not None else None
    ctx = context_for(A)
    ctx.call("cpab_prep_trels", dim, nC, A.data_ptr(), int(sign), float(dt),
             sA.data_ptr() if sA is not None else 0, T.data_ptr())


def gpu_expm(As, Ts_vectorized, p=12):
    """cpab/gpu/expm/expm_affine_by_series_2D.py:246-258 (and _3D.py:228-246):
    dim is inferred from the row length (6 -> 2D, 12 -> 3D)."""
    A = dev_tensor(As, "float64")
    T = dev_tensor(Ts_vectorized, "float64")
    if T.ndim != 2 or T.shape[1] not in (6, 12):
        raise ValueError(tuple(T.shape))
    A2 = A.reshape(A.shape[0], -1)
    if tuple(A2.shape) != tuple(T.shape):
        raise ValueError(tuple(A.shape), tuple(T.shape))
    dim = 2 if T.shape[1] == 6 else 3
    ctx = context_for(A)
    ctx.call("cpab_expm", dim, int(T.shape[0]), A2.data_ptr(), T.data_ptr(), int(p))
