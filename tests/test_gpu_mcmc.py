"""Device Metropolis registration vs the CPU reference chain (oracle + numpy
Philox), graph replay vs launches, and one chain alone vs inside a batch."""
import sys

import numpy as np
import pytest

from mcmc_ref import log_likelihood, run_chain

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def problem(gpu):
    from cpabdiffeo_amd.registration import RegistrationProblem
    return RegistrationProblem(n=64, seed=3, nTimeSteps=50, nStepsODE=10)


def test_target_image_matches_oracle(problem, orc):
    say("registration target")
    p = problem
    Av = p.B @ p.theta_star
    sA, T = orc.prep(2, Av.reshape(p.space.nC, -1), -1, p.params["dt"])
    inv = orc.calc_T(p.space.tessellation.geom(), p.space.tessellation.dense_grid_img(), T, sA, p.params["dt"],
                     p.params["nTimeSteps"], p.params["nStepsODEsolver"])
    warped = orc.remap2d(inv, p.src)
    noise = p.dst - warped
    assert abs(noise.std() - 0.02) < 0.002  # the rest is the N(0, 0.02^2) noise


@pytest.fixture(scope="module")
def problem60(gpu):
    """60 x 60 pixels: not a multiple of 8, so pixels go to waves in row order
    (and the cells are 15 pixels wide: non-power-of-two lookups)."""
    from cpabdiffeo_amd.registration import RegistrationProblem
    return RegistrationProblem(n=60, seed=4, nTimeSteps=50, nStepsODE=10)


@pytest.mark.parametrize("spec,size,ppw", [(1, 64, 0), (4, 64, 0), (1, 60, 0), (4, 60, 0), (4, 64, 128)])
def test_chain_matches_cpu_reference(problem, problem60, orc, gpu, spec, size, ppw):
    """ppw 128: two pixels per lane, pixel blocks dispatched in work order."""
    from cpabdiffeo_amd.registration import DeviceMetropolis
    say("chain vs CPU reference, 2 chains x 12 steps, spec_depth", spec, "image", size, "pixels per wave", ppw)
    p = problem if size == 64 else problem60
    nsteps = 12
    mh = DeviceMetropolis(p, nChains=2, scale=0.05, seed=11, trace_len=nsteps, spec_depth=spec, pts_per_wave=ppw)
    theta0 = np.zeros(mh.d)
    mh.set_theta(theta0)
    ll0 = log_likelihood(orc, p.space, p.B, theta0, p.src, p.dst, p.sigma, p.params)
    assert abs(mh.ll.cpu().numpy()[0] - ll0) <= 1e-9 * abs(ll0)
    mh.run(nsteps, use_graph=False)
    trace = mh.trace.cpu().numpy()
    theta = mh.theta.cpu().numpy()
    nacc = mh.n_accepted.cpu().numpy()
    for ch in range(2):
        ref = run_chain(orc, p.space, p.B, mh.Lprop, theta0, p.src, p.dst, p.sigma, p.params, nsteps, seed=11,
                        chain=ch)
        acc = [r[0] for r in ref]
        assert nacc[ch] == sum(acc), (ch, nacc[ch], acc)
        np.testing.assert_allclose(trace[ch], [r[2] for r in ref], rtol=1e-9)
        np.testing.assert_allclose(theta[ch], ref[-1][1], rtol=1e-12, atol=1e-14)
    assert 0 < nacc.sum() < 2 * nsteps  # both branches of the test ran


@pytest.mark.parametrize("spec", [1, 4])
def test_chain_with_log_prior_matches_cpu_reference(problem, orc, gpu, spec):
    """The log-prior term of Metropolis.step (Metropolis.py:64-70:
    log_ratio += wlp (lp' - lp), lp = -0.5 theta^T Sigma_CPA^-1 theta as
    ScaleDependentLogPrior computes it): the device chain with the prior
    equals the CPU reference chain with it, and the term changes the chain
    (a weight large enough to matter at this proposal scale)."""
    from cpabdiffeo_amd.registration import DeviceMetropolis
    say("chain with log-prior vs CPU reference, spec_depth", spec)
    p = problem
    nsteps, wlp = 16, 2.0
    runs = {}
    for use_prior in (True, False):
        mh = DeviceMetropolis(p, nChains=2, scale=0.05, seed=21, trace_len=nsteps, spec_depth=spec,
                              use_prior=use_prior, wlp=wlp)
        mh.set_theta(np.zeros(mh.d))
        mh.run(nsteps, use_graph=False)
        runs[use_prior] = (mh.trace.cpu().numpy(), mh.theta.cpu().numpy(), mh.n_accepted.cpu().numpy(),
                           mh.prior_prec, mh.Lprop)
        mh.close()
    trace, theta, nacc, J, Lp = runs[True]
    for ch in range(2):
        ref = run_chain(orc, p.space, p.B, Lp, np.zeros(len(theta[ch])), p.src, p.dst, p.sigma, p.params, nsteps,
                        seed=21, chain=ch, J=J, wlp=wlp)
        acc = [r[0] for r in ref]
        assert nacc[ch] == sum(acc), (ch, nacc[ch], acc)
        np.testing.assert_allclose(trace[ch], [r[2] for r in ref], rtol=1e-9)
        np.testing.assert_allclose(theta[ch], ref[-1][1], rtol=1e-12, atol=1e-14)
    assert 0 < nacc.sum() < 2 * nsteps
    # the prior term is active: the chains differ from the ones without it
    assert not np.array_equal(runs[False][0], trace)


def test_graph_replay_equals_launches_and_batch_equals_alone(problem, gpu):
    import torch
    from cpabdiffeo_amd.registration import DeviceMetropolis
    say("graph replay and batching")
    p = problem
    res = []
    for nChains, graph in ((1, False), (1, True), (6, True)):
        mh = DeviceMetropolis(p, nChains=nChains, scale=0.05, seed=5, trace_len=20)
        mh.set_theta(np.zeros(mh.d))
        mh.run(20, use_graph=graph)
        torch.cuda.synchronize()
        res.append((mh.theta.cpu().numpy(), mh.trace.cpu().numpy(), mh.n_accepted.cpu().numpy()))
        mh.close()
    for k in (1, 2):
        assert np.array_equal(res[k][0][0], res[0][0][0])
        assert np.array_equal(res[k][1][0], res[0][1][0])
        assert res[k][2][0] == res[0][2][0]
    # the other chains of the batch are different chains
    assert not np.array_equal(res[2][0][1], res[2][0][0])


def test_fallback_counter_matches_calc_T(problem, orc, gpu):
    from cpabdiffeo_amd.registration import DeviceMetropolis
    say("fallback counter")
    p = problem
    mh = DeviceMetropolis(p, nChains=1, scale=0.05, seed=2, count_fallbacks=True, integrate_corner=True)
    theta = p.theta_star.copy()
    mh.set_theta(theta)
    got = int(mh.n_fallback.cpu().numpy()[0])
    Av = p.B @ theta
    sA, T = orc.prep(2, Av.reshape(p.space.nC, -1), -1, p.params["dt"])
    _, nfb, _ = orc.calc_T(p.space.tessellation.geom(), p.space.tessellation.dense_grid_img(), T, sA,
                           p.params["dt"], p.params["nTimeSteps"], p.params["nStepsODEsolver"], probe=True)
    assert abs(got - int(nfb.sum())) <= 0.01 * nfb.sum() + 2


@pytest.mark.parametrize("spec,ppw", [(1, 0), (8, 0), (1, 256)])
def test_corner_shortcut_is_the_same_chain(problem, gpu, spec, ppw):
    """The domain-corner pixel left out where the bound on its path fixes its
    sample to src(0, 0): theta, the ll trace, the accept count and the step
    counter are bit-identical to the chain that integrates it, and the
    shortcut ran (the corner's rejected Trel proposals, ~1 per step, are
    missing from the fallback count)."""
    import torch
    from cpabdiffeo_amd.registration import DeviceMetropolis
    say("corner shortcut", spec, ppw)
    p = problem
    out = []
    for integ in (True, False):
        mh = DeviceMetropolis(p, nChains=1, scale=0.05, seed=11, trace_len=60, count_fallbacks=True,
                              spec_depth=spec, pts_per_wave=ppw, integrate_corner=integ)
        mh.set_theta(p.theta_star.copy())
        mh.run(60)
        torch.cuda.synchronize()
        out.append([x.cpu().numpy().copy() for x in (mh.theta, mh.trace, mh.n_accepted, mh.step_counter,
                                                       mh.n_fallback)])
        mh.close()
    a, b = out
    for i in range(4):
        assert np.array_equal(a[i], b[i]), i
    assert 0 < a[2][0] < 60
    missing = int(a[4][0] - b[4][0])
    say("corner fallbacks left out", missing)
    assert 30 <= missing <= 61 * p.params["nTimeSteps"]  # 60 steps and the start's evaluation


@pytest.mark.parametrize("nChains,spec,graph", [(1, 8, False), (1, 8, True), (3, 5, True), (1, 0, True)])
def test_speculation_is_the_same_chain(problem, gpu, nChains, spec, graph):
    """K proposals per launch on the all-accepted path, decided in order:
    the chain (theta, trace, accept count, fallback count) is bit-identical
    to one step per launch, across rejections and run boundaries."""
    import torch
    from cpabdiffeo_amd.registration import DeviceMetropolis
    say("speculation", nChains, spec, graph)
    p = problem
    out = []
    for k in (1, spec):
        mh = DeviceMetropolis(p, nChains=nChains, scale=0.05, seed=9, trace_len=60, count_fallbacks=True,
                              spec_depth=k)
        if k == spec and spec != 1:
            assert mh.spec_depth > 1
        mh.set_theta(np.zeros(mh.d))
        for n in (23, 1, 36):  # run boundaries inside launches
            mh.run(n, use_graph=graph)
        torch.cuda.synchronize()
        out.append([x.cpu().numpy().copy() for x in (mh.theta, mh.trace, mh.n_accepted, mh.n_fallback,
                                                       mh.step_counter, mh.stats)])
        depth = mh.spec_depth
        mh.close()
    a, b = out
    for i in range(5):
        assert np.array_equal(a[i], b[i]), i
    assert (a[4] == 60).all()
    st = b[5]
    assert (st[:, 3] == 60).all()  # steps decided
    assert (st[:, 0] >= 60).all()  # proposals evaluated in full
    npx = p.n * p.n
    assert (st[:, 4] >= st[:, 0] * npx).all()  # pixels integrated: at least those of the full proposals
    assert (st[:, 4] <= st[:, 2] * depth * npx).all()  # and never more than the slots launched
    assert 0 < a[2].min() and a[2].max() < 60  # accepts and rejects both happened


def test_changing_the_speculation_depth_keeps_the_chain(problem, gpu):
    import torch
    from cpabdiffeo_amd.registration import DeviceMetropolis
    say("speculation depth changed between runs")
    p = problem
    out = []
    for plan in ([(1, 60)], [(0, 11), (5, 7), (32, 20), (1, 2), (16, 20)]):
        mh = DeviceMetropolis(p, nChains=1, scale=0.05, seed=4, trace_len=60, count_fallbacks=True,
                              spec_depth=plan[0][0])
        mh.set_theta(np.zeros(mh.d))
        for K, n in plan:
            if K:
                mh.set_spec_depth(K)
            mh.run(n)
        torch.cuda.synchronize()
        out.append([x.cpu().numpy().copy() for x in (mh.theta, mh.trace, mh.n_accepted, mh.n_fallback, mh.step_counter)])
        mh.close()
    for i in range(5):
        assert np.array_equal(out[0][i], out[1][i]), i


def test_chain_alone_equals_chain_in_a_batch_at_equal_pixels_per_wave(problem, gpu):
    import torch
    from cpabdiffeo_amd.registration import DeviceMetropolis
    say("alone vs batch at 256 pixels per wave")
    p = problem
    res = []
    for nChains in (1, 5):
        mh = DeviceMetropolis(p, nChains=nChains, scale=0.05, seed=8, trace_len=16, pts_per_wave=256)
        mh.set_theta(np.zeros(mh.d))
        mh.run(16)
        torch.cuda.synchronize()
        res.append((mh.theta.cpu().numpy()[0].copy(), mh.trace.cpu().numpy()[0].copy()))
        mh.close()
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])


def test_many_chains_one_step_per_launch_match_cpu_reference(orc, gpu):
    """256 chains of 128 x 128 pixels, one step per launch: the library's
    many-chain configuration (one-wave blocks, four pixels per lane, blocks
    in work order, the wave-uniform fast lookup).  Two of the chains against
    the CPU reference chain."""
    from cpabdiffeo_amd.registration import DeviceMetropolis, RegistrationProblem
    p = RegistrationProblem(n=128, seed=6, nTimeSteps=50, nStepsODE=10)
    nsteps = 6
    say("256 chains x 128x128 x", nsteps, "steps vs CPU reference")
    mh = DeviceMetropolis(p, nChains=256, scale=0.05, seed=13, trace_len=nsteps, spec_depth=1)
    theta0 = np.zeros(mh.d)
    mh.set_theta(theta0)
    mh.run(nsteps)
    trace = mh.trace.cpu().numpy()
    theta = mh.theta.cpu().numpy()
    nacc = mh.n_accepted.cpu().numpy()
    mh.close()
    for ch in (0, 255):
        ref = run_chain(orc, p.space, p.B, mh.Lprop, theta0, p.src, p.dst, p.sigma, p.params, nsteps, seed=13,
                        chain=ch)
        assert nacc[ch] == sum(r[0] for r in ref), ch
        np.testing.assert_allclose(trace[ch], [r[2] for r in ref], rtol=1e-9)
        np.testing.assert_allclose(theta[ch], ref[-1][1], rtol=1e-12, atol=1e-14)
    assert 0 < nacc.sum() < 256 * nsteps
