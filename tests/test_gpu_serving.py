"""The serving loop: every result equals the same request served alone
(single calc_T call + remap), the HBM budget holds, and iterations mix
requests (continuous batching)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def kinds(gpu):
    from cpabdiffeo_amd import serving
    return serving.default_kinds()


def test_results_equal_alone_and_in_the_mix(kinds):
    import torch
    from cpabdiffeo_amd import serving
    reqs = serving.make_mix(40, kinds, rate=4000.0, seed=1, n1d=(1e3, 3e4), sides=(32, 64, 128))
    srv = serving.Server(kinds, max_points=1 << 16, tile=4096, params=dict(dt=0.02, nTimeSteps=50, nStepsODEsolver=10))
    st = srv.run(reqs, keep=True)
    say("served", st)
    assert st["requests"] == len(reqs)
    assert st["iterations_with_several_requests"] > 0
    for r in reqs:
        alone = srv.alone(r)
        torch.cuda.synchronize()
        assert torch.equal(alone, r.result), (r.id, r.kind)


def test_budget_is_respected(kinds):
    from cpabdiffeo_amd import serving
    reqs = serving.make_mix(30, kinds, rate=1e6, seed=2, n1d=(1e3, 1e4), sides=(32, 64))
    budget = 12 << 20
    srv = serving.Server(kinds, budget_bytes=budget, max_points=1 << 15, tile=2048,
                         params=dict(dt=0.02, nTimeSteps=20, nStepsODEsolver=10))
    assert max(srv.bytes_of(r) for r in reqs) <= budget
    st = srv.run(reqs, realtime=False)
    say("budget run", st)
    assert st["requests"] == len(reqs)
    assert st["peak_bytes"] <= budget


def test_theta2avees_is_batch_independent(kinds):
    """Avees = B theta on the device: the same bits for a theta alone and
    inside a stack, and each row the index-order sum."""
    import torch
    from cpabdiffeo_amd import serving
    srv = serving.Server(kinds, n_volumes=1)
    k = kinds["3d"]
    rng = np.random.default_rng(0)
    th = rng.standard_normal((5, k.d))
    stack = srv.avees("3d", torch.from_numpy(th).cuda()).cpu().numpy()
    one = srv.avees("3d", torch.from_numpy(th[2:3]).cuda()).cpu().numpy()
    assert np.array_equal(stack[2], one[0])
    B = k.space.B
    for r in (0, 17, B.shape[0] - 1):
        acc = 0.0
        for j in range(k.d):
            acc += B[r, j] * th[4, j]
        assert stack[4, r] == acc


def test_autotune_keeps_the_lowest_p99_schedule(kinds):
    from cpabdiffeo_amd import serving
    reqs = serving.make_mix(30, kinds, rate=3000.0, seed=4, n1d=(1e3, 2e4), sides=(32, 64))
    srv = serving.Server(kinds, n_volumes=1, params=dict(dt=0.02, nTimeSteps=20, nStepsODEsolver=10))
    res = srv.autotune(reqs, candidates=((1024, 1024, 1 << 16), (4096, 256, 1 << 15)))
    best = min(res, key=lambda k: res[k]["p99_ms"])
    assert (srv.tile, srv.sched.block, srv.max_points) == best
    assert all(v["requests"] == len(reqs) for v in res.values())
