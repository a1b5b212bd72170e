"""Device-resident MCMC image registration (the config-2 / config-3b path).

Mirrors the reference's registration loop: Metropolis
(cpab/cpaNd/inference/Metropolis.py:14-151) driving a Gaussian random-walk
Proposal (cpab/cpaNd/inference/Proposal.py:15-64) and the Gaussian SSD
likelihood of a warped image
(cpab/cpa2d/model/transformations/register/_ScaleDependentLogLikelihoodGaussian.py:34-109).
The reference runs that loop on the host with one blocking device->host copy
per step; here theta, the log-likelihoods, the counters and the random
numbers stay on the GPU and many chains advance in the same launches.

Choices, each switchable to the reference's behaviour:
 * proposal covariance: the full scale^2 * Sigma_CPA (``diagonal=False``) or
   its diagonal only, as the reference does (Proposal.py:58-64);
 * acceptance: the Metropolis rule (``inverted_acceptance=False``) or the
   reference's test log(u) > log_ratio (Metropolis.py:77-82);
 * ``use_ave_ll`` (default True, as the reference): the log-likelihood is
   divided by the pixel count (Metropolis.py:31-32, 58-59).
"""
import ctypes
import time

import numpy as np

from . import configs
from ._lib import Context, check, lib
from .gpuarray import require_gpu


class RegDesc(ctypes.Structure):
    """cpab_reg2d_desc of include/cpab_hip.h."""
    _fields_ = [("nChains", ctypes.c_int), ("d", ctypes.c_int), ("tess", ctypes.c_int), ("nC", ctypes.c_int),
                ("nC0", ctypes.c_int), ("nC1", ctypes.c_int), ("inc_x", ctypes.c_double),
                ("inc_y", ctypes.c_double), ("H", ctypes.c_int), ("W", ctypes.c_int),
                ("sigma", ctypes.c_double), ("dt", ctypes.c_double), ("nTimeSteps", ctypes.c_int),
                ("nStepsODE", ctypes.c_int), ("use_ave_ll", ctypes.c_int), ("inverted_acceptance", ctypes.c_int),
                ("seed", ctypes.c_uint64), ("B", ctypes.c_void_p), ("Lprop", ctypes.c_void_p),
                ("src", ctypes.c_void_p), ("dst", ctypes.c_void_p), ("theta", ctypes.c_void_p),
                ("ll", ctypes.c_void_p), ("n_accepted", ctypes.c_void_p), ("step", ctypes.c_void_p),
                ("trace", ctypes.c_void_p), ("trace_len", ctypes.c_int), ("n_fallback", ctypes.c_void_p),
                ("spec_depth", ctypes.c_int), ("stats", ctypes.c_void_p), ("pts_per_wave", ctypes.c_int),
                ("wlp", ctypes.c_double), ("prior_prec", ctypes.c_void_p), ("integrate_corner", ctypes.c_int)]


def pixel_basis(space_unit, n):
    """Basis acting on the pixel domain [0, n]^2: the unit-domain basis with
    the translation rows scaled by n (same velocity fields, pixel units)."""
    B = space_unit.B.copy().reshape(space_unit.nC, 2, 3, space_unit.d)
    B[:, :, 2, :] *= n
    return B.reshape(space_unit.D, space_unit.d)


class RegistrationProblem:
    """Source/target images, the pixel-domain CPA space and the prior.

    The target is src warped by T^{-1} of theta* (a prior draw) plus
    N(0, noise^2) pixel noise, computed on the GPU (calc_T_inv + remap_fwd)."""

    def __init__(self, n=256, seed=0, nCs=(4, 4), noise=0.02, sigma=0.02, n_blobs=40,
                 nTimeSteps=100, nStepsODE=10, t=1.0, length_scale=0.1, variance=1.0):
        torch = require_gpu()
        self.n = n
        self.sigma = sigma
        self.space_unit = configs.unit_space(nCs, "I")
        self.space = configs.pixel_space(n, nCs, "I")
        self.cov = self.space_unit.se_prior_cov(length_scale, variance)
        self.B = pixel_basis(self.space_unit, n)
        self.params = dict(dt=t / nTimeSteps, nTimeSteps=nTimeSteps, nStepsODEsolver=nStepsODE)
        self.src = configs.blob_image(n, n_blobs, seed=seed)
        rng = np.random.default_rng(seed + 1)
        self.theta_star = self.space_unit.sample_theta(self.cov, rng)
        grid = self.space.tessellation.dense_grid_img()
        self.space.update_pat(self.B @ self.theta_star)
        pts = torch.from_numpy(grid).cuda()
        inv = torch.zeros_like(pts)
        self.space.calc_T_inv(pts=pts, out=inv, **self.params)
        warped = torch.zeros((n, n), dtype=torch.float32, device="cuda")
        ctx = Context.get(pts.device.index or 0)
        ctx.bind_stream(torch.cuda.current_stream().cuda_stream)
        src_d = torch.from_numpy(self.src).cuda()
        ctx.call("cpab_remap2d", inv.data_ptr(), src_d.data_ptr(), warped.data_ptr(), n, n)
        self.dst = (warped.cpu().numpy() + rng.normal(0, noise, (n, n))).astype(np.float32)


class DeviceMetropolis:
    """nChains Metropolis chains of one RegistrationProblem, all on the GPU.

    The API follows Metropolis (set_theta / step / run / record); theta has
    shape (nChains, d).  spec_depth: proposals evaluated per chain and launch
    (1: one step per launch; K > 1: speculative prefetch of the next K steps,
    the same chain bit for bit; 0: chosen by the library).  pts_per_wave:
    pixels per wave (0: by the launch size); chains run with the same value
    give the same bits alone and in a batch.  use_prior: add the log-prior
    ratio wlp (lp' - lp), lp = -0.5 theta^T Sigma_CPA^-1 theta, to the
    acceptance test as Register does (Metropolis.py:64-70,
    Register.py:224-248, wlp = 1e-4 there); prior_prec overrides the
    precision matrix.  integrate_corner: integrate the domain-corner pixel
    even when a bound on its path fixes its sample to src(0, 0) (the chain is
    the same either way; only the fallback counts and stats include it)."""

    def __init__(self, problem, nChains=1, scale=0.01, diagonal=False, use_ave_ll=True,
                 inverted_acceptance=False, seed=0, trace_len=0, count_fallbacks=False, spec_depth=1,
                 pts_per_wave=0, use_prior=False, wlp=1e-4, prior_prec=None, integrate_corner=False):
        torch = require_gpu()
        self.problem = problem
        self.nChains = nChains
        sp = problem.space
        d = problem.space_unit.d
        self.d = d
        cov = problem.cov
        if diagonal:
            L = np.diag(scale * np.sqrt(np.clip(np.diag(cov), 0, None)))
        else:
            C = scale ** 2 * cov
            L = np.linalg.cholesky(C + 1e-12 * np.trace(C) / d * np.eye(d))
        self.Lprop = L
        dev = "cuda"
        f64 = torch.float64
        self._B = torch.from_numpy(np.ascontiguousarray(problem.B)).to(dev)
        self._L = torch.from_numpy(np.ascontiguousarray(L)).to(dev)
        self._src = torch.from_numpy(problem.src).to(dev)
        self._dst = torch.from_numpy(problem.dst).to(dev)
        self.theta = torch.zeros((nChains, d), dtype=f64, device=dev)
        self.ll = torch.zeros(nChains, dtype=f64, device=dev)
        self.n_accepted = torch.zeros(nChains, dtype=torch.int64, device=dev)
        self.step_counter = torch.zeros(nChains, dtype=torch.int64, device=dev)
        self.trace_len = trace_len
        self.trace = torch.zeros((nChains, max(trace_len, 1)), dtype=f64, device=dev)
        self.n_fallback = torch.zeros(nChains, dtype=torch.int64, device=dev)
        # proposals evaluated in full, rejected Trel proposals of every pixel integrated,
        # launches, steps decided, pixels integrated
        self.stats = torch.zeros((nChains, 5), dtype=torch.int64, device=dev)
        self.wlp = float(wlp) if use_prior else 0.0
        self.prior_prec = None
        if use_prior:
            J = np.linalg.inv(cov) if prior_prec is None else np.asarray(prior_prec, dtype=np.float64)
            self.prior_prec = J
            self._J = torch.from_numpy(np.ascontiguousarray(J)).to(dev)
        p = problem.params
        T = sp.tessellation
        desc = RegDesc(nChains, d, T.type, sp.nC, int(T.nCs[0]), int(T.nCs[1]), float(T.incs[0]),
                       float(T.incs[1]), problem.n, problem.n, float(problem.sigma), float(p["dt"]),
                       int(p["nTimeSteps"]), int(p["nStepsODEsolver"]), int(use_ave_ll), int(inverted_acceptance),
                       int(seed), self._B.data_ptr(), self._L.data_ptr(), self._src.data_ptr(),
                       self._dst.data_ptr(), self.theta.data_ptr(), self.ll.data_ptr(), self.n_accepted.data_ptr(),
                       self.step_counter.data_ptr(), self.trace.data_ptr() if trace_len else None, int(trace_len),
                       self.n_fallback.data_ptr() if count_fallbacks else None, int(spec_depth),
                       self.stats.data_ptr(), int(pts_per_wave), self.wlp,
                       self._J.data_ptr() if use_prior else None, int(bool(integrate_corner)))
        self._ctx = Context.get(0)
        self._ctx.bind_stream(torch.cuda.current_stream().cuda_stream)
        h = ctypes.c_void_p()
        check(lib().cpab_reg2d_create(self._ctx.handle, ctypes.byref(desc), ctypes.byref(h)), self._ctx.handle)
        self._h = h
        k = ctypes.c_int()
        check(lib().cpab_reg2d_spec_depth(h, ctypes.byref(k)), self._ctx.handle)
        self.spec_depth = k.value
        self.nSteps = 0
        self.record = {"runs": []}

    def _bind(self):
        torch = require_gpu()
        self._ctx.bind_stream(torch.cuda.current_stream().cuda_stream)

    def set_theta(self, theta):
        """Metropolis.set_theta (Metropolis.py:25-40): set every chain's theta
        and evaluate its log-likelihood."""
        torch = require_gpu()
        th = np.broadcast_to(np.asarray(theta, dtype=np.float64), (self.nChains, self.d))
        self.theta.copy_(torch.from_numpy(np.ascontiguousarray(th)))
        self._bind()
        check(lib().cpab_reg2d_init(self._h), self._ctx.handle)

    def step(self):
        self.run(1)

    def run(self, N, use_graph=True):
        """N steps of every chain (Metropolis.run, :104-151).  With spec_depth 1
        this only enqueues; with speculation it returns once the steps are done."""
        self._bind()
        t0 = time.perf_counter()
        check(lib().cpab_reg2d_run(self._h, int(N), 1 if use_graph else 0), self._ctx.handle)
        self.nSteps += N
        self.record["runs"].append({"N": N, "enqueue_s": time.perf_counter() - t0})

    def set_spec_depth(self, K):
        """Proposals per chain and launch from now on (the chain is unchanged)."""
        self._bind()
        check(lib().cpab_reg2d_set_spec_depth(self._h, int(K)), self._ctx.handle)
        self.spec_depth = int(K)

    def launch_counts(self):
        """(step launches replayed from captured graphs, launched directly) so far."""
        a, b = ctypes.c_int64(), ctypes.c_int64()
        check(lib().cpab_reg2d_launch_counts(self._h, ctypes.byref(a), ctypes.byref(b)), self._ctx.handle)
        return a.value, b.value

    def autotune_spec(self, candidates=(8, 16, 24, 32, 40, 48, 56, 64), steps=192, use_graph=True):
        """Time `steps` Metropolis steps of these chains at each candidate
        speculation depth (the steps are real steps of the chains) and keep
        the fastest.  Returns {depth: steps per second}."""
        torch = require_gpu()
        rates = {}
        for K in candidates:
            try:
                self.set_spec_depth(K)
            except ValueError:
                continue  # beyond what this chain set was sized for
            self.run(16 * K, use_graph=use_graph)  # capture (a graph is kept per depth) and warm
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            self.run(steps, use_graph=use_graph)
            torch.cuda.synchronize()
            rates[K] = steps / (time.perf_counter() - t0)
        best = max(rates, key=rates.get)
        self.set_spec_depth(best)
        return rates

    def last_proposal(self):
        """(theta' the next step will evaluate, ll of the last evaluated proposal)."""
        torch = require_gpu()
        tp = torch.zeros((self.nChains, self.d), dtype=torch.float64, device="cuda")
        lp = torch.zeros(self.nChains, dtype=torch.float64, device="cuda")
        check(lib().cpab_reg2d_last_proposal(self._h, tp.data_ptr(), lp.data_ptr()), self._ctx.handle)
        return tp, lp

    def accept_ratio(self):
        return self.n_accepted.double().cpu().numpy() * 100.0 / max(self.nSteps, 1)

    def close(self):
        if getattr(self, "_h", None):
            lib().cpab_reg2d_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HostMetropolis:
    """The plain path, one GPU call sequence per proposal, structured as the
    reference's loop: Metropolis.step (cpab/cpaNd/inference/Metropolis.py:41-102)
    on the host, the proposal theta' = theta + L z and Avees = B theta' in
    numpy (Proposal.py:29-64, _CpaSpace.py:366-376), update_pat, then
    calc_T_inv, remap and the SSD reduction on the GPU and a blocking copy of
    the log-likelihood to the host (the reference's ``.get()``,
    register/gpu/gaussian.py:15-46).  Random numbers come from numpy, so its
    chain is not the device chain; it is the baseline the batched device
    paths are measured against."""

    def __init__(self, problem, scale=0.01, use_ave_ll=True, inverted_acceptance=False, seed=0):
        torch = require_gpu()
        self.problem = problem
        sp = problem.space
        self.d = problem.space_unit.d
        C = scale ** 2 * problem.cov
        self.Lprop = np.linalg.cholesky(C + 1e-12 * np.trace(C) / self.d * np.eye(self.d))
        self.use_ave_ll = use_ave_ll
        self.inverted = inverted_acceptance
        self.rng = np.random.default_rng(seed)
        n = problem.n
        self.pts = torch.from_numpy(sp.tessellation.dense_grid_img()).cuda()
        self.inv = torch.zeros_like(self.pts)
        self.src = torch.from_numpy(problem.src).cuda()
        self.dst = torch.from_numpy(problem.dst).cuda()
        self.warped = torch.zeros((n, n), dtype=torch.float32, device="cuda")
        self.ll_dev = torch.zeros(1, dtype=torch.float64, device="cuda")
        self.ctx = Context.get(0)
        self.nSteps = 0
        self.n_accepted = 0

    def log_likelihood(self, theta):
        torch = require_gpu()
        p = self.problem
        p.space.update_pat(p.B @ theta)
        p.space.calc_T_inv(pts=self.pts, out=self.inv, **p.params)
        self.ctx.bind_stream(torch.cuda.current_stream().cuda_stream)
        n = p.n
        self.ctx.call("cpab_remap2d", self.inv.data_ptr(), self.src.data_ptr(), self.warped.data_ptr(), n, n)
        self.ctx.call("cpab_ssd_ll", self.warped.data_ptr(), self.dst.data_ptr(), n * n, float(p.sigma),
                      self.ll_dev.data_ptr())
        ll = float(self.ll_dev.cpu().numpy()[0])
        return ll / (n * n) if self.use_ave_ll else ll

    def set_theta(self, theta):
        self.theta = np.array(theta, dtype=np.float64)
        self.ll = self.log_likelihood(self.theta)

    def step(self):
        prop = self.theta + self.Lprop @ self.rng.standard_normal(self.d)
        llp = self.log_likelihood(prop)
        lr = llp - self.ll
        if lr > 0:
            acc = True
        else:
            lu = np.log(self.rng.random())
            acc = (lu > lr) if self.inverted else (lu < lr)
        if acc:
            self.theta, self.ll = prop, llp
            self.n_accepted += 1
        self.nSteps += 1
        return acc

    def run(self, N):
        for _ in range(N):
            self.step()
