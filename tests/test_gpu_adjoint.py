"""The adjoint kernels (calc_T_simple VJP and the fused SSD loss + gradient)
against the oracle's forward sensitivity (the reference's own derivative,
calc_2Dto2D.cu:202-319), against central finite differences, and through
the torch.autograd.Functions."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


def basmats(B, nC):
    return np.ascontiguousarray(B.T.reshape(B.shape[1], nC, 2, 3))


@pytest.fixture(scope="module")
def pix(gpu):
    from cpabdiffeo_amd import configs
    from cpabdiffeo_amd.registration import pixel_basis
    n = 32
    su = configs.unit_space((4, 4), "I")
    sp = configs.pixel_space(n, (4, 4), "I")
    cov = su.se_prior_cov(0.1, 1.0)
    return dict(n=n, su=su, sp=sp, cov=cov, B=pixel_basis(su, n),
                params=dict(dt=0.05, nTimeSteps=20, nStepsODEsolver=10))


@pytest.mark.parametrize("case", ["I-pixel", "II-unit"])
def test_vjp_equals_forward_sensitivity(pix, orc, case):
    import torch
    from cpabdiffeo_amd import adjoint, configs
    say("vjp vs sensitivity", case)
    rng = np.random.default_rng(1)
    if case == "I-pixel":
        sp, B, cov, ext = pix["sp"], pix["B"], pix["cov"], pix["n"]
        theta = pix["su"].sample_theta(cov, rng)
    else:
        sp = configs.unit_space((4, 4), "II")
        B, ext = sp.B, 1.0
        theta = sp.sample_theta(sp.se_prior_cov(0.2, 1.0), rng)
    params = pix["params"]
    sign = -1
    nC = sp.nC
    As = (sign * (B @ theta)).reshape(nC, 6)
    N = 700
    pts = rng.uniform(0, ext, (N, 2))
    lam = rng.standard_normal((N, 2))
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    out = torch.empty((1, N, 2), dtype=torch.float64, device="cuda")
    gradA = adjoint.calc_T_simple_vjp(sp, t(pts), t(lam), t(As), params, pos_out=out)
    g_adj = sign * (gradA.cpu().numpy().reshape(-1) @ B)
    geom = sp.tessellation.geom()
    pos_ref, G = orc.calc_grad_theta(geom, pts, As, sign * basmats(B, nC), params["dt"], params["nTimeSteps"],
                                     params["nStepsODEsolver"])
    assert np.array_equal(out.cpu().numpy()[0], pos_ref)  # the same RK2, bit for bit
    g_ref = np.einsum("ni,nij->j", lam, G)
    err = np.abs(g_adj - g_ref).max() / np.abs(g_ref).max()
    say("max rel err", err)
    assert err < 1e-10


def _pairs(pix, orc, nInst=2, noise=0.01):
    from cpabdiffeo_amd import configs
    rng = np.random.default_rng(7)
    n, sp, B, params = pix["n"], pix["sp"], pix["B"], pix["params"]
    geom = sp.tessellation.geom()
    grid = sp.tessellation.dense_grid_img()
    src, dst, star = [], [], []
    for i in range(nInst):
        s = configs.blob_image(n, 20, seed=10 + i).astype(np.float32)
        th = pix["su"].sample_theta(pix["cov"], rng)
        As = (-(B @ th)).reshape(sp.nC, 6)
        end = orc.calc_T_simple(geom, grid, As, params["dt"], params["nTimeSteps"], params["nStepsODEsolver"])
        w, _, _ = orc.bilinear64(s, end)
        dst.append((w.reshape(n, n) + noise * rng.standard_normal((n, n))).astype(np.float32))
        src.append(s)
        star.append(th)
    return np.stack(src), np.stack(dst), np.stack(star)


def _oracle_loss_grad(orc, pix, src, dst, theta, sign=-1):
    sp, B, params = pix["sp"], pix["B"], pix["params"]
    geom = sp.tessellation.geom()
    grid = sp.tessellation.dense_grid_img()
    As = (sign * (B @ theta)).reshape(sp.nC, 6)
    pos, G = orc.calc_grad_theta(geom, grid, As, sign * basmats(B, sp.nC), params["dt"], params["nTimeSteps"],
                                 params["nStepsODEsolver"])
    w, gx, gy = orc.bilinear64(src, pos)
    r = w - dst.reshape(-1).astype(np.float64)
    lam = np.stack([r * gx, r * gy], axis=1)
    return 0.5 * np.sum(r * r), np.einsum("ni,nij->j", lam, G)


def test_ssd_loss_gradient_vs_reference_and_fd(pix, orc):
    import torch
    from cpabdiffeo_amd import adjoint
    say("SSD adjoint vs reference gradient and central differences")
    src, dst, star = _pairs(pix, orc)
    sp, B, params = pix["sp"], pix["B"], pix["params"]
    theta = 0.5 * star
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    Bt = t(B)

    def gpu(theta_np):
        As = (-(theta_np @ B.T)).reshape(len(theta_np), sp.nC, 6)
        loss, gradA = adjoint.ssd_loss_and_grad(sp, t(As), t(src), t(dst), params)
        return loss.cpu().numpy(), -(gradA.cpu().numpy().reshape(len(theta_np), -1) @ B)

    loss, g = gpu(theta)
    for i in range(len(theta)):
        L_ref, g_ref = _oracle_loss_grad(orc, pix, src[i], dst[i], theta[i])
        assert abs(loss[i] - L_ref) <= 1e-12 * abs(L_ref)
        err = np.abs(g[i] - g_ref).max() / np.abs(g_ref).max()
        say("instance", i, "loss", loss[i], "grad rel err vs reference", err)
        assert err < 1e-8
    rng = np.random.default_rng(3)
    h = 1e-6
    for k in range(3):
        v = rng.standard_normal(theta.shape)
        lp, _ = gpu(theta + h * v)
        lm, _ = gpu(theta - h * v)
        fd = (lp - lm) / (2 * h)
        an = np.sum(g * v, axis=1)
        rel = np.abs(fd - an) / np.maximum(np.abs(an), 1e-3 * np.linalg.norm(g, axis=1) * np.linalg.norm(v, axis=1))
        say("direction", k, "fd", fd, "adjoint", an, "rel", rel)
        assert (rel < 1e-4).all()


def test_autograd_functions(pix, orc):
    import torch
    from cpabdiffeo_amd import adjoint
    say("autograd functions")
    src, dst, star = _pairs(pix, orc)
    sp, B, params = pix["sp"], pix["B"], pix["params"]
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    Bt = t(B)
    theta = t(0.3 * star).requires_grad_(True)
    loss = adjoint.warp_ssd(theta, Bt, t(src), t(dst), sp, params, sign=-1)
    (loss * t(np.array([1.0, 2.0]))).sum().backward()
    As = (-(0.3 * star @ B.T)).reshape(2, sp.nC, 6)
    l2, gradA = adjoint.ssd_loss_and_grad(sp, t(As), t(src), t(dst), params)
    want = -(gradA.cpu().numpy().reshape(2, -1) @ B) * np.array([[1.0], [2.0]])
    np.testing.assert_allclose(loss.detach().cpu().numpy(), l2.cpu().numpy(), rtol=1e-13)
    np.testing.assert_allclose(theta.grad.cpu().numpy(), want, rtol=1e-11, atol=1e-14)
    # the transform: d/dtheta sum(lam * T(pts)) = the VJP
    rng = np.random.default_rng(5)
    pts = t(rng.uniform(0, pix["n"], (300, 2)))
    lam = rng.standard_normal((300, 2))
    th = t(star[0]).requires_grad_(True)
    out = adjoint.transform_simple(th, pts, Bt, sp, params, sign=1)
    geom = sp.tessellation.geom()
    # the tables exactly as the Function formed them (GEMV on the device)
    As1 = (th.detach().reshape(1, -1) @ Bt.t().contiguous()).cpu().numpy().reshape(sp.nC, 6)
    ref_end = orc.calc_T_simple(geom, pts.cpu().numpy(), As1, params["dt"], params["nTimeSteps"],
                                params["nStepsODEsolver"])
    assert np.array_equal(out.detach().cpu().numpy(), ref_end)
    (out * t(lam)).sum().backward()
    _, G = orc.calc_grad_theta(geom, pts.cpu().numpy(), As1, basmats(B, sp.nC), params["dt"],
                               params["nTimeSteps"], params["nStepsODEsolver"])
    g_ref = np.einsum("ni,nij->j", lam, G)
    err = np.abs(th.grad.cpu().numpy() - g_ref).max() / np.abs(g_ref).max()
    assert err < 1e-10


def test_gradient_descent_reduces_the_loss(pix, orc):
    import torch
    from cpabdiffeo_amd import adjoint
    say("gradient descent")
    src, dst, star = _pairs(pix, orc, noise=0.0)
    sp, B, params = pix["sp"], pix["B"], pix["params"]
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    Bt, S, Dd = t(B), t(src), t(dst)
    theta = torch.zeros((2, len(star[0])), dtype=torch.float64, device="cuda", requires_grad=True)
    opt = torch.optim.SGD([theta], lr=2e-3)
    hist = []
    for it in range(15):
        opt.zero_grad()
        loss = adjoint.warp_ssd(theta, Bt, S, Dd, sp, params, sign=-1)
        loss.sum().backward()
        opt.step()
        hist.append(loss.detach().cpu().numpy())
    hist = np.array(hist)
    say("loss", hist[0], "->", hist[-1])
    assert (np.diff(hist, axis=0) < 0).all()  # every step descends
    assert (hist[-1] < 0.97 * hist[0]).all()
