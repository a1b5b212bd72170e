"""The drop-in surface, used the way a reference user would: TransformWrapper
(cpab/cpa2d/TransformWrapper.py) with update_pat_from_Avees, calc_T_fwd /
calc_T_inv, calc_v, calc_cell_idx, calc_trajectory and remap_fwd on
CpuGpuArrays, against the oracle and the reference demo's invariant that a
trajectory ends where calc_T ends (cpab/gpu/demo.py:77-81)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def tw_setup(gpu):
    from cpabdiffeo_amd import TransformWrapper
    from cpabdiffeo_amd.transform_wrapper import Bunch
    tw = TransformWrapper(nRows=96, nCols=128, tess="I", zero_v_across_bdry=[True, True], valid_outside=False,
                          base=[4, 4])
    tw.params_flow_int = Bunch(dt=0.01, nTimeSteps=100, nStepsODEsolver=10)
    sp = tw.ms.L_cpa_space[0]
    rng = np.random.default_rng(11)
    theta = 30.0 * sp.sample_theta(sp.se_prior_cov(0.1, 1.0), rng)
    Avees = sp.theta2Avees(theta)
    tw.update_pat_from_Avees(Avees, level=0)
    pts = np.stack([rng.uniform(0, 128, 3000), rng.uniform(0, 96, 3000)], axis=1)
    return tw, sp, np.asarray(Avees).reshape(sp.nC, 6), pts


def test_calc_T_fwd_inv_and_trajectory(tw_setup, orc):
    from cpabdiffeo_amd import CpuGpuArray
    tw, sp, As, pts = tw_setup
    p = tw.params_flow_int
    src = CpuGpuArray(pts)
    fwd = CpuGpuArray.zeros_like(src)
    tw.calc_T_fwd(src, fwd, level=0)
    fwd.gpu2cpu()
    sA, T = orc.prep(2, As, 1, p.dt)
    want, _, md = orc.calc_T(sp.tessellation.geom(), pts, T, sA, p.dt, p.nTimeSteps, p.nStepsODEsolver, probe=True)
    far = md >= 1e-12
    err = np.abs(fwd.cpu - want).max(axis=1)
    say("calc_T_fwd max err", err[far].max(), "near-facet points", int((~far).sum()), "moved", np.abs(want - pts).max())
    assert err[far].max() <= 1e-10
    assert np.abs(want - pts).max() > 1.0  # the transform is not close to the identity
    back = CpuGpuArray.zeros_like(src)
    tw.calc_T_inv(fwd, back, level=0)
    back.gpu2cpu()
    assert np.abs(back.cpu - pts).max() < 1e-3  # T^{-1}(T(x)) = x up to the discretization
    traj = tw.calc_trajectory(pts=src, mysign=1, level=0)
    traj.gpu2cpu()
    assert traj.cpu.shape == (1 + p.nTimeSteps, len(pts), 2)
    assert np.array_equal(traj.cpu[0], pts)
    assert np.array_equal(traj.cpu[-1], fwd.cpu)  # bitwise, the demo's 1e-6 tightened


def test_calc_v_and_cell_idx(tw_setup, orc):
    from cpabdiffeo_amd import CpuGpuArray
    tw, sp, As, pts = tw_setup
    src = CpuGpuArray(pts)
    v = CpuGpuArray.zeros_like(src)
    tw.calc_v(level=0, pts=src, v=v)
    v.gpu2cpu()
    assert np.array_equal(v.cpu, orc.calc_v(sp.tessellation.geom(), pts, As))
    idx = CpuGpuArray.zeros((len(pts),), dtype=np.int32)
    tw.calc_cell_idx(src, idx, level=0)
    idx.gpu2cpu()
    assert np.array_equal(idx.cpu, orc.calc_cell_idx(sp.tessellation.geom(), pts))


def test_remap_fwd_of_the_dense_grid(tw_setup, orc):
    from cpabdiffeo_amd import CpuGpuArray, configs
    tw, sp, As, pts = tw_setup
    img = configs.blob_image(128, 30, seed=4)[:96].copy()
    grid = tw.pts_src_dense
    inv = CpuGpuArray.zeros_like(grid)
    tw.calc_T_inv(grid, inv, level=0)
    inv.gpu2cpu()
    src_img = CpuGpuArray(img.astype(np.float32))
    out = CpuGpuArray.zeros((96, 128), dtype=np.float32)
    tw.remap_fwd(inv, src_img, out)
    out.gpu2cpu()
    assert np.array_equal(out.cpu, orc.remap2d(inv.cpu, img.astype(np.float32)))
    assert not np.array_equal(out.cpu, img)  # it warped
