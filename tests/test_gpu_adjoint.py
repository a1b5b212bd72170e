"""The adjoint kernels -- of calc_T's map (Trel steps + RK2 redos, the map
of calc_T_fwd / calc_T_inv) and of calc_T_simple's RK2 map; VJP and the
fused SSD loss + gradient -- against the oracle's forward sensitivities
(calc_T's, and the reference's own solveODE2, calc_2Dto2D.cu:202-319),
against central finite differences, run to run (bit for bit), and through
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
    loss = adjoint.warp_ssd(theta, Bt, t(src), t(dst), sp, params, sign=-1, simple=True)
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


@pytest.mark.parametrize("simple", [False, True])
def test_gradient_descent_reduces_the_loss(pix, orc, simple):
    import torch
    from cpabdiffeo_amd import adjoint
    say("gradient descent, simple map:", simple)
    src, dst, star = _pairs(pix, orc, noise=0.0)
    sp, B, params = pix["sp"], pix["B"], pix["params"]
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    Bt, S, Dd = t(B), t(src), t(dst)
    theta = torch.zeros((2, len(star[0])), dtype=torch.float64, device="cuda", requires_grad=True)
    opt = torch.optim.SGD([theta], lr=2e-3)
    hist = []
    for it in range(15):
        opt.zero_grad()
        loss = adjoint.warp_ssd(theta, Bt, S, Dd, sp, params, sign=-1, simple=simple)
        loss.sum().backward()
        opt.step()
        hist.append(loss.detach().cpu().numpy())
    hist = np.array(hist)
    say("loss", hist[0], "->", hist[-1])
    assert (np.diff(hist, axis=0) < 0).all()  # every step descends
    assert (hist[-1] < 0.97 * hist[0]).all()


# ---------------------------------------------------------------- calc_T's map


def _dev_tables(torch, theta, B, sign, nC, dt):
    """(signed tables, Trels) as the library forms them for calc_T."""
    from cpabdiffeo_amd import adjoint
    sA, T = adjoint._tables(torch, torch.from_numpy(np.atleast_2d(theta)).cuda(), torch.from_numpy(B).cuda(), sign,
                            nC, dt)
    return sA, T


@pytest.mark.parametrize("case,nS", [("I-pixel", 10), ("II-unit", 10), ("I-pixel", 4), ("I-pixel", 13)])
def test_calc_T_vjp_equals_forward_sensitivity(pix, orc, case, nS):
    """cpab_calc_T_vjp vs the oracle's forward sensitivity of calc_T (pinned
    by finite differences on the CPU) contracted with the cotangents, with
    the same tables: the same end points bit for bit, gradients to 1e-9.
    nStepsODE 4 runs part of the kernel's 10 substep registers, 13 the form
    with 16."""
    import torch
    from cpabdiffeo_amd import adjoint, configs
    say("calc_T vjp vs sensitivity", case, "nS", nS)
    rng = np.random.default_rng(11)
    if case == "I-pixel":
        sp, B, cov, ext = pix["sp"], pix["B"], pix["cov"], pix["n"]
        theta = pix["su"].sample_theta(cov, rng)
    else:
        sp = configs.unit_space((4, 4), "II")
        B, ext = sp.B, 1.0
        theta = sp.sample_theta(sp.se_prior_cov(0.2, 1.0), rng)
    params = dict(pix["params"], nStepsODEsolver=nS)
    nC = sp.nC
    N = 700
    pts = rng.uniform(0, ext, (N, 2))
    lam = rng.standard_normal((N, 2))
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    geom = sp.tessellation.geom()
    for sign in (1, -1):
        sA, T = _dev_tables(torch, theta, B, sign, nC, params["dt"])
        out = torch.empty((1, N, 2), dtype=torch.float64, device="cuda")
        gradA = adjoint.calc_T_vjp(sp, t(pts), t(lam), sA[0], T[0], params, pos_out=out)
        g_adj = sign * (gradA.cpu().numpy().reshape(-1) @ B)
        sA_h, T_h = sA[0].cpu().numpy(), T[0].cpu().numpy()
        Bs = sign * basmats(B, nC)
        dTs = orc.trel_sensitivity(2, sA_h, Bs, params["dt"])
        pos_ref, G = orc.calc_grad_theta_T(geom, pts, T_h, sA_h, Bs, dTs, params["dt"], params["nTimeSteps"],
                                           params["nStepsODEsolver"])
        _, nfb, _ = orc.calc_T(geom, pts, T_h, sA_h, params["dt"], params["nTimeSteps"], params["nStepsODEsolver"],
                               probe=True)
        assert np.array_equal(out.cpu().numpy()[0], pos_ref)  # the same calc_T, bit for bit
        assert nfb.sum() > 0 and (nfb < params["nTimeSteps"]).all()  # both kinds of step ran
        g_ref = np.einsum("ni,nij->j", lam, G)
        err = np.abs(g_adj - g_ref).max() / np.abs(g_ref).max()
        say("sign", sign, "max rel err", err, "redone steps", int(nfb.sum()))
        assert err < 1e-9


def _oracle_loss_grad_T(orc, pix, src, dst, sA, T, sign=-1):
    """Loss and dloss/dtheta of calc_T's map from the oracle's forward
    sensitivity, given the tables."""
    sp, B, params = pix["sp"], pix["B"], pix["params"]
    geom = sp.tessellation.geom()
    grid = sp.tessellation.dense_grid_img()
    Bs = sign * basmats(B, sp.nC)
    dTs = orc.trel_sensitivity(2, sA, Bs, params["dt"])
    pos, G = orc.calc_grad_theta_T(geom, grid, T, sA, Bs, dTs, params["dt"], params["nTimeSteps"],
                                   params["nStepsODEsolver"])
    w, gx, gy = orc.bilinear64(src, pos)
    r = w - dst.reshape(-1).astype(np.float64)
    lam = np.stack([r * gx, r * gy], axis=1)
    return 0.5 * np.sum(r * r), np.einsum("ni,nij->j", lam, G)


def test_ssd_calc_T_gradient_vs_reference_fd_and_rerun(pix, orc):
    """The fused SSD adjoint of calc_T's map: loss and gradient against the
    oracle (1e-12 / 1e-8), central differences (h = 1e-6) at 1e-4 on every
    instance whose lookups are the same at theta +- h v (signatures equal:
    the map is smooth between them; the others are counted), and the same
    bits on a second run."""
    import torch
    from cpabdiffeo_amd import adjoint
    say("SSD adjoint of calc_T vs reference gradient, central differences, rerun")
    src, dst, star = _pairs(pix, orc, nInst=4)
    sp, B, params = pix["sp"], pix["B"], pix["params"]
    theta = 0.5 * star
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    n = pix["n"]

    def gpu(theta_np, sig=None):
        sA, T = _dev_tables(torch, theta_np, B, -1, sp.nC, params["dt"])
        loss, gradA = adjoint.ssd_loss_and_grad_T(sp, sA, T, t(src), t(dst), params, sig=sig)
        return loss.cpu().numpy(), -(gradA.cpu().numpy().reshape(len(theta_np), -1) @ B), sA, T, gradA

    loss, g, sA, T, gradA = gpu(theta)
    for i in range(len(theta)):
        L_ref, g_ref = _oracle_loss_grad_T(orc, pix, src[i], dst[i], sA[i].cpu().numpy(), T[i].cpu().numpy())
        assert abs(loss[i] - L_ref) <= 1e-12 * abs(L_ref)
        err = np.abs(g[i] - g_ref).max() / np.abs(g_ref).max()
        say("instance", i, "loss", loss[i], "grad rel err vs reference", err)
        assert err < 1e-8
    _, _, _, _, gradA2 = gpu(theta)
    assert torch.equal(gradA, gradA2)  # bit-reproducible
    rng = np.random.default_rng(3)
    h = 1e-6
    clean = 0
    for k in range(3):
        v = rng.standard_normal(theta.shape)
        sp_, sm_ = (torch.empty((len(theta), n * n), dtype=torch.int32, device="cuda") for _ in range(2))
        lp = gpu(theta + h * v, sp_)[0]
        lm = gpu(theta - h * v, sm_)[0]
        same = (sp_ == sm_).all(dim=1).cpu().numpy()
        fd = (lp - lm) / (2 * h)
        an = np.sum(g * v, axis=1)
        rel = np.abs(fd - an) / np.abs(an)
        say("direction", k, "same lookups", same, "rel", rel)
        assert (rel[same] < 1e-4).all()
        clean += int(same.sum())
    assert clean >= 6  # most instances are smooth over +-h


def test_transform_autograd_calc_T(pix, orc):
    """d/dtheta sum(lam * calc_T(pts)) through the Transform Function equals
    the oracle's calc_T sensitivity; forward values bit-identical to the
    oracle's calc_T with the same tables."""
    import torch
    from cpabdiffeo_amd import adjoint
    say("Transform autograd (calc_T map)")
    sp, B, params = pix["sp"], pix["B"], pix["params"]
    rng = np.random.default_rng(12)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    star = pix["su"].sample_theta(pix["cov"], rng)
    pts = t(rng.uniform(0, pix["n"], (400, 2)))
    lam = rng.standard_normal((400, 2))
    geom = sp.tessellation.geom()
    for sign in (1, -1):
        th = t(star).requires_grad_(True)
        out = adjoint.transform(th, pts, t(B), sp, params, sign=sign)
        sA, T = _dev_tables(torch, star, B, sign, sp.nC, params["dt"])
        sA_h, T_h = sA[0].cpu().numpy(), T[0].cpu().numpy()
        ref_end = orc.calc_T(geom, pts.cpu().numpy(), T_h, sA_h, params["dt"], params["nTimeSteps"],
                             params["nStepsODEsolver"])
        assert np.array_equal(out.detach().cpu().numpy(), ref_end)
        (out * t(lam)).sum().backward()
        Bs = sign * basmats(B, sp.nC)
        dTs = orc.trel_sensitivity(2, sA_h, Bs, params["dt"])
        _, G = orc.calc_grad_theta_T(geom, pts.cpu().numpy(), T_h, sA_h, Bs, dTs, params["dt"],
                                     params["nTimeSteps"], params["nStepsODEsolver"])
        g_ref = np.einsum("ni,nij->j", lam, G)
        err = np.abs(th.grad.cpu().numpy() - g_ref).max() / np.abs(g_ref).max()
        say("sign", sign, "rel err", err)
        assert err < 1e-9


def test_simple_adjoint_is_bit_reproducible(pix, orc):
    import torch
    from cpabdiffeo_amd import adjoint
    src, dst, star = _pairs(pix, orc, nInst=3)
    sp, B, params = pix["sp"], pix["B"], pix["params"]
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    As = t((-(0.5 * star @ B.T)).reshape(3, sp.nC, 6))
    r = [adjoint.ssd_loss_and_grad(sp, As, t(src), t(dst), params) for _ in range(3)]
    assert all(torch.equal(r[0][1], x[1]) and torch.equal(r[0][0], x[0]) for x in r[1:])


def test_fd_check_over_same_lookup_pixels(gpu):
    """The full-size gradient check's procedure at test size: central
    differences over the pixels whose lookups are the same at theta +- h v
    agree with the adjoint over the same pixels to 1e-4 on every pair."""
    import torch
    from cpabdiffeo_amd import adjoint
    say("FD over same-lookup pixels")
    prob = adjoint.PairsProblem(n=32, nInst=6, seed=2, nTimeSteps=50)
    rng = np.random.default_rng(4)
    theta = torch.from_numpy(0.5 * prob.theta_star).cuda()
    v = torch.from_numpy(rng.standard_normal(theta.shape)).cuda()
    fd, an, left = adjoint.fd_check_same_lookups(prob, theta, v)
    rel = ((fd - an).abs() / an.abs()).cpu().numpy()
    say("rel", rel, "pixels left out", left.cpu().numpy())
    assert (rel < 1e-4).all()
    assert int(left.max()) < prob.n * prob.n // 4
