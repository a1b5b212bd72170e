"""GPU parity of the entry points the 2D tests leave open, through the C ABI
and the drop-in wrappers, against the CPU oracle:

 * the standalone gpu_expm (cpab/gpu/expm/expm_affine_by_series_2D.py:246-258,
   expm_affine_by_series_3D.py:228-246);
 * calc_T_simple and calc_trajectory in 1D and 3D (cpab/gpu/calc_1Dto1D.cu:552,648;
   calc_3Dto3D.cu:420,516), bit-identical given the same tables, and the
   trajectory's last row equal to calc_T bitwise (cpab/gpu/demo.py:77-81);
 * the TransformWrapper1D / TransformWrapper3D surfaces
   (cpab/cpa1d/TransformWrapper.py:21-103, cpab/cpa3d/TransformWrapper.py:22-297).
"""
import sys

import numpy as np
import pytest

from test_gpu_parity import dev, host, random_field, space_for

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("dim", [2, 3])
def test_gpu_expm_standalone(gpu, orc, dim):
    """gpu_expm of an already scaled table: the 3D series and the 2D series
    branch (singular linear part) bit-identical; the 2D closed form (exp,
    cosh/sinh, cos/sin of the GPU libm) within 1e-13."""
    torch = gpu
    from cpabdiffeo_amd.calcs import gpu_expm
    say("gpu_expm", dim)
    rng = np.random.default_rng(40 + dim)
    ne = dim * (dim + 1)
    A = rng.normal(size=(700, ne)) * 0.05
    A[:20] = 0.0
    if dim == 2:
        A[20:120, 3:5] = A[20:120, 0:2] * 0.5   # det = 0 exactly (scaling by 0.5 is exact): the series branch
        A[120:130] = [0.03, 0.0, 0.02, 0.0, 0.03, -0.01]  # repeated eigenvalue
    want = orc.expm(dim, A, 12)
    T = torch.zeros(A.shape, dtype=torch.float64, device="cuda")
    gpu_expm(dev(torch, A), T, p=12)
    got = host(T)
    if dim == 3:
        assert np.array_equal(got, want)
    else:
        assert np.array_equal(got[:120], want[:120])
        assert np.abs(got - want).max() <= 1e-13
    # the reference's 3D layout passes (nC, 3, 4) tables: same result
    if dim == 3:
        T2 = torch.zeros(A.shape, dtype=torch.float64, device="cuda")
        gpu_expm(dev(torch, A.reshape(-1, 3, 4)), T2, p=12)
        assert torch.equal(T, T2)


ND = [
    ("1D-50", [0], [1], (50,), "II"),
    ("1D-8", [0], [1], (8,), "II"),
    ("3D-I-4x4x4", [0, 0, 0], [1, 1, 1], (4, 4, 4), "I"),
    ("3D-I-2x2x2-px64", [0, 0, 0], [64, 64, 64], (2, 2, 2), "I"),
    ("3D-II-2x3x4", [0, 0, 0], [1, 1, 1], (2, 3, 4), "II"),
]


@pytest.mark.parametrize("name,lo,hi,nCs,tess", ND)
def test_simple_and_trajectory_1d_3d(gpu, orc, name, lo, hi, nCs, tess):
    torch = gpu
    say("simple + trajectory", name)
    sp = space_for(lo, hi, nCs, tess)
    g = sp.tessellation.geom()
    Av = random_field(sp, 41)
    rng = np.random.default_rng(42)
    pts = rng.uniform(0, 1, (4000, sp.dim_domain)) * sp.XMAXS
    dt, nT, nS = 0.02, 50, 10
    calcs = sp._gpu_calcs
    for sign in (1, -1):
        sA, T = orc.prep(sp.dim_domain, Av.reshape(sp.nC, -1), sign, dt)
        want = orc.calc_trajectory(g, pts, T, sA, dt, nT, nS)
        traj = calcs.calc_trajectory(None, None, dev(torch, T), dev(torch, sA), dev(torch, pts), dt, nT, nS,
                                     sp.dim_domain, sp.dim_range, sp.nCs, sp.incs)
        tr = host(traj)
        assert np.array_equal(tr, want), (name, sign, np.abs(tr - want).max())
        out = torch.zeros(pts.shape, dtype=torch.float64, device="cuda")
        calcs.calc_T(None, None, dev(torch, T), dev(torch, sA), dev(torch, pts), dt, nT, nS, out,
                     sp.dim_domain, sp.dim_range, sp.nCs, sp.incs)
        assert np.array_equal(tr[-1], host(out))
        ws = orc.calc_T_simple(g, pts, sA, dt, nT, nS)
        calcs.calc_T_simple(None, None, dev(torch, sA), dev(torch, pts), dt, nT, nS, out, sp.dim_domain,
                            sp.dim_range, sp.nCs, sp.incs)
        assert np.array_equal(host(out), ws), (name, sign)
        assert np.abs(ws - pts).max() > 1e-3 * float(np.max(sp.XMAXS))  # the field moves the points


def _sample_avees(sp, seed, scale):
    rng = np.random.default_rng(seed)
    theta = scale * sp.sample_theta(sp.se_prior_cov(0.1, 1.0), rng)
    return np.asarray(sp.theta2Avees(theta)).reshape(sp.nC, -1)


def _check_wrapper(tw, sp, As, pts, orc, moved_at_least):
    """calc_T_fwd / calc_T_inv against the oracle (1e-10 away from facets),
    inverse after forward, trajectory end == calc_T bitwise, calc_v and
    calc_cell_idx bitwise, all through the wrapper on CpuGpuArrays."""
    from cpabdiffeo_amd import CpuGpuArray
    p = tw.params_flow_int
    g = sp.tessellation.geom()
    src = CpuGpuArray(pts)
    fwd = CpuGpuArray.zeros_like(src)
    tw.calc_T_fwd(src, fwd, level=0)
    fwd.gpu2cpu()
    for sign, got in ((1, fwd.cpu), (-1, None)):
        if got is None:
            inv = CpuGpuArray.zeros_like(src)
            tw.calc_T_inv(src, inv, level=0)
            inv.gpu2cpu()
            got = inv.cpu
        sA, T = orc.prep(sp.dim_domain, As, sign, p.dt)
        want, _, md = orc.calc_T(g, pts, T, sA, p.dt, p.nTimeSteps, p.nStepsODEsolver, probe=True)
        far = md >= 1e-12
        err = np.abs(got - want).max(axis=1)
        say("sign", sign, "max err", float(err[far].max()), "near-facet", int((~far).sum()),
            "moved", float(np.abs(want - pts).max()))
        assert err[far].max() <= 1e-10
        assert np.abs(want - pts).max() > moved_at_least
    back = CpuGpuArray.zeros_like(src)
    tw.calc_T_inv(fwd, back, level=0)
    back.gpu2cpu()
    assert np.abs(back.cpu - pts).max() < 1e-3 * float(np.max(sp.XMAXS))
    traj = tw.calc_trajectory(pts=src, mysign=1, level=0)
    traj.gpu2cpu()
    assert traj.cpu.shape == (1 + p.nTimeSteps, len(pts), sp.dim_domain)
    assert np.array_equal(traj.cpu[0], pts)
    assert np.array_equal(traj.cpu[-1], fwd.cpu)
    v = CpuGpuArray.zeros_like(src)
    tw.calc_v(level=0, pts=src, v=v)
    v.gpu2cpu()
    assert np.array_equal(v.cpu, orc.calc_v(g, pts, As))
    idx = CpuGpuArray.zeros((len(pts),), dtype=np.int32)
    tw.calc_cell_idx(src, idx, level=0)
    idx.gpu2cpu()
    assert np.array_equal(idx.cpu, orc.calc_cell_idx(g, pts))
    return fwd


def test_transform_wrapper_1d(gpu, orc):
    """TransformWrapper1D (cpab/cpa1d/TransformWrapper.py:21-103): 50
    intervals on [0, 100] (a non-power-of-two increment)."""
    from cpabdiffeo_amd import TransformWrapper1D
    from cpabdiffeo_amd.transform_wrapper import Bunch
    say("TransformWrapper1D")
    tw = TransformWrapper1D(nCols=100, base=(50,), zero_v_across_bdry=(True,))
    tw.params_flow_int = Bunch(dt=0.01, nTimeSteps=100, nStepsODEsolver=10)
    sp = tw.ms.L_cpa_space[0]
    As = _sample_avees(sp, 43, 40.0)
    tw.update_pat_from_Avees(As.reshape(-1), level=0)
    pts = np.random.default_rng(44).uniform(0, 100, (5000, 1))
    _check_wrapper(tw, sp, As, pts, orc, 0.5)


@pytest.mark.parametrize("tess", ["I", "II"])
def test_transform_wrapper_3d(gpu, orc, tess):
    """TransformWrapper3D (cpab/cpa3d/TransformWrapper.py:22-297) on a
    24 x 32 x 16 voxel domain, plus remap_fwd of a volume (trilinear) on the
    calc_T_inv points of the dense voxel grid, bit-identical to the oracle."""
    import torch
    from cpabdiffeo_amd import CpuGpuArray, TransformWrapper3D, configs
    from cpabdiffeo_amd.transform_wrapper import Bunch
    say("TransformWrapper3D", tess)
    nR, nCo, nS = 24, 32, 16
    tw = TransformWrapper3D(nRows=nR, nCols=nCo, nSlices=nS, base=(2, 2, 2), tess=tess,
                            zero_v_across_bdry=(True, True, True))
    tw.params_flow_int = Bunch(dt=0.01, nTimeSteps=100, nStepsODEsolver=10)
    sp = tw.ms.L_cpa_space[0]
    As = _sample_avees(sp, 45, 20.0)
    tw.update_pat_from_Avees(As.reshape(-1), level=0)
    rng = np.random.default_rng(46)
    pts = rng.uniform(0, 1, (4000, 3)) * np.array([nCo, nR, nS], dtype=np.float64)
    _check_wrapper(tw, sp, As, pts, orc, 0.5)
    zz, yy, xx = np.mgrid[0:nS, 0:nR, 0:nCo]
    grid = np.stack([xx.ravel(), yy.ravel(), zz.ravel()], 1).astype(np.float64)
    inv = CpuGpuArray.zeros_like(CpuGpuArray(grid))
    tw.calc_T_inv(CpuGpuArray(grid), inv, level=0)
    inv.gpu2cpu()
    vol = configs.blob_image(32, 20, seed=5, dims=3)[:nS, :nR, :nCo].astype(np.float32).copy()
    out = CpuGpuArray.zeros((nS, nR, nCo), dtype=np.float32)
    tw.remap_fwd(inv, CpuGpuArray(vol), out)
    out.gpu2cpu()
    assert np.array_equal(out.cpu, orc.remap3d(inv.cpu, vol))
    assert not np.array_equal(out.cpu, vol)  # it warped
    torch.cuda.synchronize()
