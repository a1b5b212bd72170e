"""ctypes wrapper of the CPU oracle (libcpab_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the cpabdiffeo_amd product path.

The C library restates the reference kernels operation by operation
(cpab/gpu/calc_{1Dto1D,2Dto2D,3Dto3D}.cu, cpab/gpu/expm/*.py,
cpab/cpaNd/_CpaCalcs.py:227-313); these wrappers take numpy arrays and the
same scalars the reference's Calcs object passes to its kernels
(cpab/gpu/Calcs.py:540-561).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "libcpab_oracle.so")
_lib = None

_d = ctypes.c_double
_i = ctypes.c_int
_i64 = ctypes.c_int64
_p = ctypes.c_void_p


def build(force=False):
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    if force or not os.path.exists(_SO):
        subprocess.check_call(["make", "-s"] + (["-B"] if force else []) + ["-C", _HERE])
    return _SO


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = ctypes.CDLL(_SO)
        geo = [_i, _i, _i, _d, _d, _d]
        L.oracle_calc_T.argtypes = [_i, _i, _p, _p, _p, _p, _d, _i, _i, _i64] + geo + [_p, _p, _i]
        L.oracle_calc_T_simple.argtypes = [_i, _i, _p, _p, _p, _d, _i, _i, _i64] + geo + [_p, _i]
        L.oracle_calc_trajectory.argtypes = [_i, _i, _p, _p, _p, _p, _d, _i, _i, _i64] + geo + [_i]
        L.oracle_calc_v.argtypes = [_i, _i, _p, _p, _p, _i64] + geo + [_i]
        L.oracle_calc_cell_idx.argtypes = [_i, _i, _p, _p, _i64] + geo + [_i]
        L.oracle_facet_dist.argtypes = [_i, _i, _p, _p, _i64] + geo
        L.oracle_calc_grad_theta.argtypes = ([_i, _i, _p, _p, _p, _p, _p, _i, _i, _d, _i, _i, _i64]
                                             + geo + [_i])
        L.oracle_calc_grad_theta_T.argtypes = ([_i, _i, _p, _p, _p, _p, _p, _p, _p, _i, _i, _d, _i, _i, _i64]
                                               + geo + [_i])
        L.oracle_prep.argtypes = [_i, _i, _p, _i, _d, _p, _p]
        L.oracle_expm.argtypes = [_i, _i, _p, _p, _i]
        L.oracle_remap2d.argtypes = [_p, _p, _p, _i, _i, _i]
        L.oracle_remap3d.argtypes = [_p, _p, _p, _i, _i, _i, _i]
        L.oracle_ssd_ll.argtypes = [_p, _p, _i64, _d]
        L.oracle_ssd_ll.restype = _d
        L.oracle_transform.argtypes = ([_i, _i, _i, _p, _i, _p, _p, _d, _i, _i, _i64] + geo + [_i])
        L.oracle_max_threads.restype = _i
        _lib = L
    return _lib


def _c(a, dtype):
    a = np.ascontiguousarray(a, dtype=dtype)
    return a, a.ctypes.data_as(_p)


def _geo(geom):
    """geom: dict/obj with dim, tess (1=I, 2=II), nCs (len<=3), incs (len<=3).

    Mirrors Calcs.parse_nCs_and_incs (cpab/gpu/Calcs.py:215-233): unused
    counts are 0, unused increments are 1.0."""
    nCs = list(geom["nCs"]) + [0] * (3 - len(geom["nCs"]))
    incs = list(geom["incs"]) + [1.0] * (3 - len(geom["incs"]))
    return [int(nCs[0]), int(nCs[1]), int(nCs[2]), float(incs[0]), float(incs[1]), float(incs[2])]


def prep(dim, As, sign, dt):
    """sign*As and expm(dt*sign*As) per cell (_CpaCalcs.py:243-313)."""
    As, pA = _c(As, np.float64)
    nC = As.shape[0]
    sA = np.empty_like(As)
    T = np.empty_like(As)
    lib().oracle_prep(dim, nC, pA, int(sign), float(dt), sA.ctypes.data_as(_p),
                      T.ctypes.data_as(_p))
    return sA, T


def expm(dim, As, p=12):
    As, pA = _c(As, np.float64)
    T = np.empty_like(As)
    lib().oracle_expm(dim, As.shape[0], pA, T.ctypes.data_as(_p), int(p))
    return T


def calc_T(geom, pos0, Trels, As, dt, nTimeSteps, nStepsODE, probe=False, nthreads=0):
    dim = geom["dim"]
    pos0, p0 = _c(pos0, np.float64)
    nPts = pos0.shape[0]
    out = np.empty_like(pos0)
    Trels, pT = _c(Trels, np.float64)
    As, pA = _c(As, np.float64)
    nfb = np.zeros(nPts, np.int32) if probe else None
    md = np.zeros(nPts, np.float64) if probe else None
    lib().oracle_calc_T(dim, geom["tess"], p0, out.ctypes.data_as(_p), pT, pA, float(dt),
                        int(nTimeSteps), int(nStepsODE), nPts, *_geo(geom),
                        nfb.ctypes.data_as(_p) if probe else None,
                        md.ctypes.data_as(_p) if probe else None, int(nthreads))
    if probe:
        return out, nfb, md
    return out


def calc_T_simple(geom, pos0, As, dt, nTimeSteps, nStepsODE, probe=False, nthreads=0):
    pos0, p0 = _c(pos0, np.float64)
    out = np.empty_like(pos0)
    As, pA = _c(As, np.float64)
    md = np.zeros(pos0.shape[0], np.float64) if probe else None
    lib().oracle_calc_T_simple(geom["dim"], geom["tess"], p0, out.ctypes.data_as(_p), pA,
                               float(dt), int(nTimeSteps), int(nStepsODE), pos0.shape[0],
                               *_geo(geom), md.ctypes.data_as(_p) if probe else None,
                               int(nthreads))
    return (out, md) if probe else out


def calc_trajectory(geom, pos0, Trels, As, dt, nTimeSteps, nStepsODE, nthreads=0):
    pos0, p0 = _c(pos0, np.float64)
    nPts, dim = pos0.shape
    traj = np.empty((int(nTimeSteps), nPts, dim), np.float64)
    Trels, pT = _c(Trels, np.float64)
    As, pA = _c(As, np.float64)
    lib().oracle_calc_trajectory(geom["dim"], geom["tess"], p0, traj.ctypes.data_as(_p), pT, pA,
                                 float(dt), int(nTimeSteps), int(nStepsODE), nPts, *_geo(geom),
                                 int(nthreads))
    return traj


def calc_v(geom, pts, As, nthreads=0):
    pts, pp = _c(pts, np.float64)
    v = np.empty_like(pts)
    As, pA = _c(As, np.float64)
    lib().oracle_calc_v(geom["dim"], geom["tess"], pp, v.ctypes.data_as(_p), pA, pts.shape[0],
                        *_geo(geom), int(nthreads))
    return v


def calc_cell_idx(geom, pts, nthreads=0):
    pts, pp = _c(pts, np.float64)
    out = np.empty(pts.shape[0], np.int32)
    lib().oracle_calc_cell_idx(geom["dim"], geom["tess"], pp, out.ctypes.data_as(_p),
                               pts.shape[0], *_geo(geom), int(nthreads))
    return out


def facet_dist(geom, pts):
    pts, pp = _c(pts, np.float64)
    out = np.empty(pts.shape[0], np.float64)
    lib().oracle_facet_dist(geom["dim"], geom["tess"], pp, out.ctypes.data_as(_p), pts.shape[0],
                            *_geo(geom))
    return out


def calc_grad_theta(geom, pos0, As, Bs, dt, nTimeSteps, nStepsODE, nthreads=0):
    """Forward sensitivity (calc_2Dto2D.cu:202-319).  Bs: (d, nC, dim, dim+1)."""
    pos0, p0 = _c(pos0, np.float64)
    nPts, dim = pos0.shape
    Bs, pB = _c(Bs, np.float64)
    d, nC = Bs.shape[0], Bs.shape[1]
    As, pA = _c(As, np.float64)
    pos = np.empty_like(pos0)
    grad = np.empty((nPts, dim, d), np.float64)
    lib().oracle_calc_grad_theta(geom["dim"], geom["tess"], p0, pos.ctypes.data_as(_p), pA, pB,
                                 grad.ctypes.data_as(_p), d, nC, float(dt), int(nTimeSteps),
                                 int(nStepsODE), nPts, *_geo(geom), int(nthreads))
    return pos, grad


def trel_sensitivity(dim, sA, Bs, dt):
    """dTrel_c / dtheta_j = the top dim rows of the Frechet derivative of the
    matrix exponential at dt * A~_c in the direction dt * B~_jc (A~, B~ the
    (dim+1) x (dim+1) affine matrices with a zero last row), computed by
    scipy.linalg.expm_frechet (Al-Mohy and Higham, SIAM J. Matrix Anal.
    Appl. 30(4), 2009), independent of any code here.  sA (nC, dim(dim+1)):
    the signed tables Trel = expm(dt sA); Bs (d, nC, dim, dim+1) = dsA /
    dtheta.  Returns dTs with Bs's layout."""
    from scipy.linalg import expm_frechet
    sA = np.asarray(sA, dtype=np.float64).reshape(-1, dim, dim + 1)
    Bs = np.asarray(Bs, dtype=np.float64)
    d, nC = Bs.shape[0], Bs.shape[1]
    out = np.zeros((d, nC, dim, dim + 1))
    X = np.zeros((dim + 1, dim + 1))
    E = np.zeros((dim + 1, dim + 1))
    for c in range(nC):
        X[:dim] = dt * sA[c]
        for j in range(d):
            E[:dim] = dt * Bs[j, c]
            out[j, c] = expm_frechet(X, E, compute_expm=False)[:dim]
    return out


def calc_grad_theta_T(geom, pos0, Trels, As, Bs, dTs, dt, nTimeSteps, nStepsODE, nthreads=0):
    """Forward sensitivity of calc_T (accepted Trel steps and RK2 redos, the
    decisions of calc_2Dto2D.cu:348-408 held).  Bs, dTs: (d, nC, dim, dim+1)
    derivatives of the signed tables and of the Trels in theta.  Returns
    (calc_T end points, grad (nPts, dim, d))."""
    pos0, p0 = _c(pos0, np.float64)
    nPts, dim = pos0.shape
    Bs, pB = _c(Bs, np.float64)
    dTs, pD = _c(dTs, np.float64)
    d, nC = Bs.shape[0], Bs.shape[1]
    As, pA = _c(As, np.float64)
    Trels, pT = _c(Trels, np.float64)
    pos = np.empty_like(pos0)
    grad = np.empty((nPts, dim, d), np.float64)
    lib().oracle_calc_grad_theta_T(geom["dim"], geom["tess"], p0, pos.ctypes.data_as(_p), pT, pA, pB, pD,
                                   grad.ctypes.data_as(_p), d, nC, float(dt), int(nTimeSteps),
                                   int(nStepsODE), nPts, *_geo(geom), int(nthreads))
    return pos, grad


def remap2d(pts, img, nthreads=0):
    img, pi = _c(img, np.float32)
    H, W = img.shape
    pts, pp = _c(pts, np.float64)
    out = np.empty_like(img)
    lib().oracle_remap2d(pp, pi, out.ctypes.data_as(_p), H, W, int(nthreads))
    return out


def remap3d(pts, vol, nthreads=0):
    vol, pv = _c(vol, np.float32)
    D, H, W = vol.shape
    pts, pp = _c(pts, np.float64)
    out = np.empty_like(vol)
    lib().oracle_remap3d(pp, pv, out.ctypes.data_as(_p), D, H, W, int(nthreads))
    return out


def ssd_ll(x, y, sigma):
    x, px = _c(x, np.float32)
    y, py = _c(y, np.float32)
    return lib().oracle_ssd_ll(px, py, x.size, float(sigma))


def transform(geom, nC, As, sign, pos0, dt, nTimeSteps, nStepsODE, nthreads=0):
    """Timed CPU unit of work: prep (sign, dt, expm) + calc_T."""
    As, pA = _c(As, np.float64)
    pos0, p0 = _c(pos0, np.float64)
    out = np.empty_like(pos0)
    lib().oracle_transform(geom["dim"], geom["tess"], int(nC), pA, int(sign), p0,
                           out.ctypes.data_as(_p), float(dt), int(nTimeSteps), int(nStepsODE),
                           pos0.shape[0], *_geo(geom), int(nthreads))
    return out


def max_threads():
    return lib().oracle_max_threads()


def bilinear64(img, pts):
    """The remap convention (OpenCV INTER_LINEAR, zero outside, a full pixel
    or more outside -> 0) evaluated in float64, with its gradient: returns
    (w, dw/dx, dw/dy) at pts (N, 2) = (x = column, y = row).  This is the
    interpolant the SSD adjoint differentiates (no reference counterpart: the
    reference's remap is the un-vendored of.gpu.remap)."""
    img = np.asarray(img, dtype=np.float64)
    H, W = img.shape
    x = np.asarray(pts[:, 0], dtype=np.float64)
    y = np.asarray(pts[:, 1], dtype=np.float64)
    with np.errstate(invalid="ignore"):
        xf, yf = np.floor(x), np.floor(y)
        ok = (xf >= -1) & (xf < W) & (yf >= -1) & (yf < H)
    xf = np.where(ok, xf, 0.0)
    yf = np.where(ok, yf, 0.0)
    fx, fy = np.where(ok, x - xf, 0.0), np.where(ok, y - yf, 0.0)
    x0, y0 = xf.astype(np.int64), yf.astype(np.int64)

    def px(yy, xx):
        inside = ok & (xx >= 0) & (xx < W) & (yy >= 0) & (yy < H)
        return np.where(inside, img[np.clip(yy, 0, H - 1), np.clip(xx, 0, W - 1)], 0.0)

    v00, v01 = px(y0, x0), px(y0, x0 + 1)
    v10, v11 = px(y0 + 1, x0), px(y0 + 1, x0 + 1)
    gx, gy = 1.0 - fx, 1.0 - fy
    a = gx * v00 + fx * v01
    b = gx * v10 + fx * v11
    w = np.where(ok, gy * a + fy * b, 0.0)
    dwdx = np.where(ok, gy * (v01 - v00) + fy * (v11 - v10), 0.0)
    dwdy = np.where(ok, b - a, 0.0)
    return w, dwdx, dwdy
