"""Gradients of the CPA transform and of the image SSD loss with respect to
theta, from the hand-written adjoint kernels (cpab_calc_T_vjp,
cpab_ssd_grad2d_calc_T, cpab_calc_T_simple_vjp, cpab_ssd_grad2d), wrapped as
torch.autograd.Functions.

Two maps are differentiated:
 * calc_T (cpab/gpu/calc_2Dto2D.cu:348-408), the map of calc_T_fwd /
   calc_T_inv: Trel proposals (Trel = expm(dt A), expm_affine_by_series_2D.py)
   kept while they stay in their cell, RK2 redos otherwise; the default;
 * calc_T_simple's RK2 map (calc_2Dto2D.cu:412-445), the map whose forward
   sensitivity the reference computes (solveODE2, calc_2Dto2D.cu:202-319;
   CpaSpace.calc_grad_theta here).
theta enters through Avees = B theta (CpaSpace.theta2Avees,
cpab/cpaNd/_CpaSpace.py:366-376) and the signed tables sign * Avees, so
dL/dtheta = sign * B^T dL/dA.

torch is plumbing here: tensors carry device memory and the autograd tape;
the derivative is computed by the library.
"""
from .calcs import context_for
from .gpuarray import require_gpu
from .tessellation import tess_code


def _geom2d(space):
    T = space.tessellation
    if space.dim_domain != 2:
        raise NotImplementedError("the adjoint kernels are 2D")
    return tess_code(space.tess), space.nC, int(T.nCs[0]), int(T.nCs[1]), float(T.incs[0]), float(T.incs[1])


def _params(params):
    return float(params["dt"]), int(params["nTimeSteps"]), int(params["nStepsODEsolver"])


def _signed_tables(torch, theta, B, sign, nC):
    Av = theta @ B.t().contiguous()
    return (float(sign) * Av).reshape(theta.shape[0], nC, 6).contiguous()


def _tables(torch, theta, B, sign, nC, dt):
    """(signed tables sign * B theta, Trels = expm(dt * signed tables)) for a
    (m, d) stack of thetas, both (m, nC, 6); the prep on the device
    (cpab_prep_trels, cpab/cpaNd/_CpaCalcs.py:227-313)."""
    Av = (theta @ B.t().contiguous()).contiguous()
    m = theta.shape[0]
    sA = torch.empty((m, nC, 6), dtype=torch.float64, device=theta.device)
    T = torch.empty_like(sA)
    context_for(Av).call("cpab_prep_trels", 2, m * nC, Av.data_ptr(), int(sign), float(dt), sA.data_ptr(),
                         T.data_ptr())
    return sA, T


def calc_T_vjp(space, pts, lam, As, Trels, params, pos_out=None):
    """gradA (nInst, nC, 6) = sum_x lam(x)^T dT(x)/dA for calc_T's map with the
    signed tables As and Trels = expm(dt As) (nInst, nC, 6); pts, lam:
    (nInst, N, 2) device float64."""
    torch = require_gpu()
    tess, nC, nC0, nC1, ix, iy = _geom2d(space)
    dt, nT, nS = _params(params)
    if As.dim() == 2:
        As, Trels = As.unsqueeze(0), Trels.unsqueeze(0)
    if pts.dim() == 2:
        pts, lam = pts.unsqueeze(0), lam.unsqueeze(0)
    nInst, N = pts.shape[0], pts.shape[1]
    if (tuple(As.shape) != (nInst, nC, 6) or tuple(Trels.shape) != (nInst, nC, 6)
            or tuple(lam.shape) != tuple(pts.shape) or pts.shape[2] != 2):
        raise ValueError((tuple(As.shape), tuple(Trels.shape), tuple(pts.shape), tuple(lam.shape)))
    gradA = torch.empty((nInst, nC, 6), dtype=torch.float64, device=pts.device)
    ctx = context_for(pts)
    ctx.call("cpab_calc_T_vjp", tess, nC, nInst, pts.contiguous().data_ptr(), lam.contiguous().data_ptr(),
             Trels.contiguous().data_ptr(), As.contiguous().data_ptr(), dt, nT, nS, int(N), nC0, nC1, ix, iy,
             pos_out.data_ptr() if pos_out is not None else None, gradA.data_ptr())
    return gradA


def calc_T_simple_vjp(space, pts, lam, As, params, pos_out=None):
    """gradA (nInst, nC, 6) = sum_x lam(x)^T dT(x)/dA for the signed tables As
    (nInst, nC, 6); pts, lam: (nInst, N, 2) device float64."""
    torch = require_gpu()
    tess, nC, nC0, nC1, ix, iy = _geom2d(space)
    dt, nT, nS = _params(params)
    if As.dim() == 2:
        As = As.unsqueeze(0)
    if pts.dim() == 2:
        pts, lam = pts.unsqueeze(0), lam.unsqueeze(0)
    nInst, N = pts.shape[0], pts.shape[1]
    if tuple(As.shape) != (nInst, nC, 6) or tuple(lam.shape) != tuple(pts.shape) or pts.shape[2] != 2:
        raise ValueError((tuple(As.shape), tuple(pts.shape), tuple(lam.shape)))
    gradA = torch.empty((nInst, nC, 6), dtype=torch.float64, device=pts.device)
    ctx = context_for(pts)
    ctx.call("cpab_calc_T_simple_vjp", tess, nC, nInst, pts.contiguous().data_ptr(), lam.contiguous().data_ptr(),
             As.contiguous().data_ptr(), dt, nT, nS, int(N), nC0, nC1, ix, iy,
             pos_out.data_ptr() if pos_out is not None else None, gradA.data_ptr())
    return gradA


def ssd_loss_and_grad(space, As, src, dst, params, pos_out=None):
    """loss (nInst,) = 0.5 sum_x (src_i(T_i(x)) - dst_i(x))^2 over the pixel
    grid and gradA (nInst, nC, 6) = dloss/dA, for signed tables As
    (nInst, nC, 6) and float32 images src, dst (nInst, H, W)."""
    torch = require_gpu()
    tess, nC, nC0, nC1, ix, iy = _geom2d(space)
    dt, nT, nS = _params(params)
    nInst, H, W = src.shape
    if tuple(dst.shape) != (nInst, H, W) or tuple(As.shape) != (nInst, nC, 6):
        raise ValueError((tuple(src.shape), tuple(dst.shape), tuple(As.shape)))
    if src.dtype != torch.float32 or dst.dtype != torch.float32:
        raise ValueError("images are float32")
    loss = torch.empty(nInst, dtype=torch.float64, device=src.device)
    gradA = torch.empty((nInst, nC, 6), dtype=torch.float64, device=src.device)
    ctx = context_for(src)
    ctx.call("cpab_ssd_grad2d", tess, nC, nInst, As.contiguous().data_ptr(), src.contiguous().data_ptr(),
             dst.contiguous().data_ptr(), int(H), int(W), dt, nT, nS, nC0, nC1, ix, iy,
             pos_out.data_ptr() if pos_out is not None else None, loss.data_ptr(), gradA.data_ptr())
    return loss, gradA


def ssd_loss_and_grad_T(space, As, Trels, src, dst, params, pos_out=None, sig=None, n_redo=None):
    """ssd_loss_and_grad for calc_T's map (Trel steps + RK2 redos) with the
    signed tables As and Trels = expm(dt As); sig (optional, int32 (nInst,
    H*W)): per pixel a hash of every cell the forward sweep looked up;
    n_redo (optional, float64 (nInst,)): the time steps redone by RK2."""
    torch = require_gpu()
    tess, nC, nC0, nC1, ix, iy = _geom2d(space)
    dt, nT, nS = _params(params)
    nInst, H, W = src.shape
    if (tuple(dst.shape) != (nInst, H, W) or tuple(As.shape) != (nInst, nC, 6)
            or tuple(Trels.shape) != (nInst, nC, 6)):
        raise ValueError((tuple(src.shape), tuple(dst.shape), tuple(As.shape), tuple(Trels.shape)))
    if src.dtype != torch.float32 or dst.dtype != torch.float32:
        raise ValueError("images are float32")
    loss = torch.empty(nInst, dtype=torch.float64, device=src.device)
    gradA = torch.empty((nInst, nC, 6), dtype=torch.float64, device=src.device)
    ctx = context_for(src)
    ctx.call("cpab_ssd_grad2d_calc_T", tess, nC, nInst, Trels.contiguous().data_ptr(), As.contiguous().data_ptr(),
             src.contiguous().data_ptr(), dst.contiguous().data_ptr(), int(H), int(W), dt, nT, nS, nC0, nC1, ix, iy,
             pos_out.data_ptr() if pos_out is not None else None, loss.data_ptr(), gradA.data_ptr(),
             sig.data_ptr() if sig is not None else None, n_redo.data_ptr() if n_redo is not None else None)
    return loss, gradA


def _function_classes():
    torch = require_gpu()

    class WarpSSD(torch.autograd.Function):
        """loss_i = 0.5 sum_x (src_i(T_{theta_i}(x)) - dst_i(x))^2 for a batch
        of image pairs on the pixel grid; theta (nInst, d), B (D, d); T the
        calc_T map (simple=False) or calc_T_simple's."""

        @staticmethod
        def forward(fctx, theta, B, src, dst, space, params, sign, simple):
            if simple:
                As = _signed_tables(torch, theta.detach(), B, sign, space.nC)
                loss, gradA = ssd_loss_and_grad(space, As, src, dst, params)
            else:
                As, T = _tables(torch, theta.detach(), B, sign, space.nC, params["dt"])
                loss, gradA = ssd_loss_and_grad_T(space, As, T, src, dst, params)
            fctx.save_for_backward(gradA, B)
            fctx.sign = float(sign)
            return loss

        @staticmethod
        def backward(fctx, gl):
            gradA, B = fctx.saved_tensors
            g = (gl[:, None] * gradA.reshape(gradA.shape[0], -1)) @ B
            return fctx.sign * g, None, None, None, None, None, None, None

    class Transform(torch.autograd.Function):
        """T_theta(pts) by calc_T (calc_T_fwd with sign 1, calc_T_inv with
        -1); its backward is the calc_T adjoint kernel.  theta (d,), pts
        (N, 2), B (D, d)."""

        @staticmethod
        def forward(fctx, theta, pts, B, space, params, sign):
            As, T = _tables(torch, theta.detach().reshape(1, -1), B, sign, space.nC, params["dt"])
            out = torch.empty_like(pts)
            tess, nC, nC0, nC1, ix, iy = _geom2d(space)
            dt, nT, nS = _params(params)
            context_for(pts).call("cpab_calc_T", 2, tess, nC, pts.contiguous().data_ptr(), out.data_ptr(),
                                  T.data_ptr(), As.data_ptr(), dt, nT, nS, int(pts.shape[0]), nC0, nC1, 0, ix, iy,
                                  1.0, None)
            fctx.save_for_backward(pts, As, T, B)
            fctx.space, fctx.params, fctx.sign = space, params, float(sign)
            return out

        @staticmethod
        def backward(fctx, gout):
            pts, As, T, B = fctx.saved_tensors
            gradA = calc_T_vjp(fctx.space, pts, gout.contiguous(), As[0], T[0], fctx.params)
            g = gradA.reshape(1, -1) @ B
            return fctx.sign * g.reshape(-1), None, None, None, None, None

    class TransformSimple(torch.autograd.Function):
        """T_theta(pts) by calc_T_simple; its backward is the adjoint kernel.
        theta (d,), pts (N, 2), B (D, d)."""

        @staticmethod
        def forward(fctx, theta, pts, B, space, params, sign):
            As = _signed_tables(torch, theta.detach().reshape(1, -1), B, sign, space.nC)
            out = torch.empty_like(pts)
            tess, nC, nC0, nC1, ix, iy = _geom2d(space)
            dt, nT, nS = _params(params)
            context_for(pts).call("cpab_calc_T_simple", 2, tess, nC, pts.contiguous().data_ptr(), out.data_ptr(),
                                  As.data_ptr(), dt, nT, nS, int(pts.shape[0]), nC0, nC1, 0, ix, iy, 1.0)
            fctx.save_for_backward(pts, As, B)
            fctx.space, fctx.params, fctx.sign = space, params, float(sign)
            return out

        @staticmethod
        def backward(fctx, gout):
            pts, As, B = fctx.saved_tensors
            gradA = calc_T_simple_vjp(fctx.space, pts, gout.contiguous(), As[0], fctx.params)
            g = gradA.reshape(1, -1) @ B
            return fctx.sign * g.reshape(-1), None, None, None, None, None

    return WarpSSD, TransformSimple, Transform


_FN = None


def warp_ssd(theta, B, src, dst, space, params, sign=-1, simple=False):
    """Differentiable batch SSD loss (see WarpSSD); sign -1 warps with the
    inverse transform, as remap_fwd does with calc_T_inv; simple: the
    calc_T_simple map instead of calc_T's."""
    global _FN
    if _FN is None:
        _FN = _function_classes()
    return _FN[0].apply(theta, B, src, dst, space, params, sign, bool(simple))


def transform(theta, pts, B, space, params, sign=1):
    """Differentiable calc_T of pts (calc_T_fwd, or calc_T_inv with sign -1;
    see Transform)."""
    global _FN
    if _FN is None:
        _FN = _function_classes()
    return _FN[2].apply(theta, pts, B, space, params, sign)


def transform_simple(theta, pts, B, space, params, sign=1):
    """Differentiable calc_T_simple of pts (see TransformSimple)."""
    global _FN
    if _FN is None:
        _FN = _function_classes()
    return _FN[1].apply(theta, pts, B, space, params, sign)


def ssd_cotangent(src, dst, pts):
    """Torch FP64 restatement, for self-checks, of what the SSD kernel
    differentiates: w = the bilinear sample of src (H, W) at pts (N, 2)
    (cpab_remap2d convention, zero a full pixel or more outside) and the
    cotangent (w - dst) * grad w of 0.5 (w - dst)^2."""
    torch = require_gpu()
    H, W = src.shape
    img = src.double()
    x, y = pts[:, 0], pts[:, 1]
    xf, yf = torch.floor(x), torch.floor(y)
    ok = (xf >= -1) & (xf < W) & (yf >= -1) & (yf < H)
    fx, fy = x - xf, y - yf
    x0, y0 = xf.long(), yf.long()

    def px(yy, xx):
        inside = ok & (xx >= 0) & (xx < W) & (yy >= 0) & (yy < H)
        return torch.where(inside, img[yy.clamp(0, H - 1), xx.clamp(0, W - 1)], torch.zeros_like(x))

    v00, v01, v10, v11 = px(y0, x0), px(y0, x0 + 1), px(y0 + 1, x0), px(y0 + 1, x0 + 1)
    a = (1 - fx) * v00 + fx * v01
    b = (1 - fx) * v10 + fx * v11
    zero = torch.zeros_like(x)
    w = torch.where(ok, (1 - fy) * a + fy * b, zero)
    gx = torch.where(ok, (1 - fy) * (v01 - v00) + fy * (v11 - v10), zero)
    gy = torch.where(ok, b - a, zero)
    r = w - dst.reshape(-1).double()
    return w, torch.stack([r * gx, r * gy], dim=1)


class PairsProblem:
    """A batch of image pairs for gradient-based registration: nInst FP32
    blob images (n x n), each target = its source warped by T^{-1} of a prior
    draw theta*_i (calc_T_inv on the pixel grid, or calc_T_simple's map with
    simple=True) plus N(0, noise^2) pixel noise; the 4 x 4 x 4-triangle
    tessellation of [0, n]^2 with zero normal boundary velocity.  The loss
    warps with the same map."""

    def __init__(self, n=128, nInst=64, seed=0, nCs=(4, 4), noise=0.02, n_blobs=40, nTimeSteps=100, nStepsODE=10,
                 t=1.0, length_scale=0.1, variance=1.0, simple=False):
        import numpy as np
        from . import configs
        from .registration import pixel_basis
        torch = require_gpu()
        self.n, self.nInst = n, nInst
        self.space_unit = configs.unit_space(nCs, "I")
        self.space = configs.pixel_space(n, nCs, "I")
        self.cov = self.space_unit.se_prior_cov(length_scale, variance)
        self.B = pixel_basis(self.space_unit, n)
        self.d = self.space_unit.d
        self.params = dict(dt=t / nTimeSteps, nTimeSteps=nTimeSteps, nStepsODEsolver=nStepsODE)
        rng = np.random.default_rng(seed)
        self.src_np = np.stack([configs.blob_image(n, n_blobs, seed=seed * 100003 + i) for i in range(nInst)])
        self.theta_star = np.stack([self.space_unit.sample_theta(self.cov, rng) for _ in range(nInst)])
        self.Bt = torch.from_numpy(np.ascontiguousarray(self.B)).cuda()
        self.src = torch.from_numpy(self.src_np).cuda()
        self.simple = bool(simple)
        # targets: src_i(T_i^{-1}(x)) + noise, the points from the same kernel
        zero = torch.zeros_like(self.src)
        pos = torch.empty((nInst, n * n, 2), dtype=torch.float64, device="cuda")
        th_star = torch.from_numpy(self.theta_star).cuda()
        if self.simple:
            As = _signed_tables(torch, th_star, self.Bt, -1, self.space.nC)
            ssd_loss_and_grad(self.space, As, self.src, zero, self.params, pos_out=pos)
        else:
            As, T = _tables(torch, th_star, self.Bt, -1, self.space.nC, self.params["dt"])
            ssd_loss_and_grad_T(self.space, As, T, self.src, zero, self.params, pos_out=pos)
        warped = torch.empty_like(self.src)
        ctx = context_for(self.src)
        for i in range(nInst):
            ctx.call("cpab_remap2d", pos[i].data_ptr(), self.src[i].data_ptr(), warped[i].data_ptr(), n, n)
        noise_t = torch.from_numpy(rng.normal(0, noise, (nInst, n, n)).astype(np.float32)).cuda()
        self.dst = (warped + noise_t).contiguous()

    def loss(self, theta):
        return warp_ssd(theta, self.Bt, self.src, self.dst, self.space, self.params, sign=-1, simple=self.simple)


def fd_check_same_lookups(problem, theta, v, h=1e-6):
    """Central differences of the calc_T-map SSD loss along v (nInst, d),
    restricted per pair to the pixels whose cell lookups, and the pixel cell
    of whose bilinear sample, are the same at theta + h v and theta - h v
    (there the loss is smooth between the two; a pixel whose accept/redo
    decision or substep cell changes jumps by the difference of a Trel step
    and its RK2 redo, and the interpolant's gradient jumps across pixel
    lines).  Returns, per pair, (the
    central difference over those pixels, the adjoint's directional
    derivative of the loss over the same pixels -- cpab_calc_T_vjp with the
    other pixels' cotangents zeroed --, the count of pixels left out)."""
    torch = require_gpu()
    sp, B, p = problem.space, problem.Bt, problem.params
    n, nI = problem.n, problem.nInst
    N = n * n

    def ends(th):
        As, T = _tables(torch, th, B, -1, sp.nC, p["dt"])
        pos = torch.empty((nI, N, 2), dtype=torch.float64, device="cuda")
        sig = torch.empty((nI, N), dtype=torch.int32, device="cuda")
        ssd_loss_and_grad_T(sp, As, T, problem.src, problem.dst, p, pos_out=pos, sig=sig)
        return As, T, pos, sig

    _, _, pp, sgp = ends(theta + h * v)
    _, _, pm, sgm = ends(theta - h * v)
    As, T, pos, _ = ends(theta)
    # the same lookups, and the same bilinear cell of the sample (the
    # interpolant's gradient jumps across pixel lines)
    same = (sgp == sgm) & (torch.floor(pp) == torch.floor(pm)).all(dim=2)
    fd = torch.empty(nI, dtype=torch.float64, device="cuda")
    lam = torch.empty((nI, N, 2), dtype=torch.float64, device="cuda")
    for i in range(nI):
        wp, _ = ssd_cotangent(problem.src[i], problem.dst[i], pp[i])
        wm, _ = ssd_cotangent(problem.src[i], problem.dst[i], pm[i])
        d = problem.dst[i].reshape(-1).double()
        lp = 0.5 * (wp - d) ** 2
        lm = 0.5 * (wm - d) ** 2
        fd[i] = ((lp - lm) * same[i]).sum() / (2 * h)
        _, l0 = ssd_cotangent(problem.src[i], problem.dst[i], pos[i])
        lam[i] = l0 * same[i].unsqueeze(1)
    grid = sp.tessellation.dense_grid_img()
    pts = torch.from_numpy(grid).cuda().unsqueeze(0).expand(nI, N, 2).contiguous()
    gA = calc_T_vjp(sp, pts, lam, As, T, p)
    an = -((gA.reshape(nI, -1) @ B) * v).sum(1)
    return fd, an, (~same).sum(1)


def gradient_descent(problem, steps, lr, theta0=None):
    """Plain gradient descent on every pair's theta (torch.optim.SGD over the
    WarpSSD Function).  Returns (theta, loss history (steps, nInst))."""
    torch = require_gpu()
    theta = (torch.zeros((problem.nInst, problem.d), dtype=torch.float64, device="cuda") if theta0 is None
             else theta0.clone()).requires_grad_(True)
    opt = torch.optim.SGD([theta], lr=lr)
    hist = []
    for _ in range(steps):
        opt.zero_grad()
        loss = problem.loss(theta)
        loss.sum().backward()
        opt.step()
        hist.append(loss.detach())
    return theta.detach(), torch.stack(hist)
