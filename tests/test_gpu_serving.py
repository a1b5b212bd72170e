"""The serving loop: every result equals the same request served alone
(single calc_T call + remap), served results match the CPU oracle at the
served mix's own sizes, the HBM budget holds, and iterations mix requests
(continuous batching)."""
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


def _request(kinds, rid, kind, n, side, seed, rng, arrival=0.0):
    from cpabdiffeo_amd import serving
    theta = kinds[kind].factor @ rng.standard_normal(kinds[kind].d)
    return serving.Request(rid, kind, arrival, n, side, seed, theta)


def test_served_results_match_the_oracle_at_mix_sizes(kinds, orc):
    """The served mix at its own sizes (1D: 50 intervals, 1e4-1e6 points; 2D
    grids up to 1024^2 on 4x4 and 8x8 x 4 triangles; 3D 64^3 volumes on
    4x4x4 x 5 tetrahedra, calc_T_inv + trilinear remap; t = 1, 100 x 10),
    against the CPU oracle (cpab/gpu/calc_1Dto1D.cu:485-548,
    calc_2Dto2D.cu:348-408, calc_3Dto3D.cu:353-416; the remap of
    cpab/cpa3d/TransformWrapper.py:254-290):
     * with the request's own device tables, points bit-identical to the
       oracle's calc_T, and the 3D warp bit-identical to the oracle's
       trilinear remap of those points;
     * end to end (the oracle's own prep of the request's Avees), points
       within 1e-10 except points whose trajectory passed within 1e-12 of a
       facet (counted);
     * every served result equals the same request served alone."""
    import torch
    from cpabdiffeo_amd import serving
    rng = np.random.default_rng(7)
    checked = [_request(kinds, 0, "1d", 1000000, 0, 11, rng),
               _request(kinds, 1, "2d4", 1024 * 1024, 1024, 12, rng),
               _request(kinds, 2, "2d8", 1024 * 1024, 1024, 13, rng, 1e-4),
               _request(kinds, 3, "3d", serving.VOL ** 3, 0, 14, rng, 2e-4),
               _request(kinds, 4, "1d", 10000, 0, 15, rng, 3e-4),
               _request(kinds, 5, "2d8", 64 * 64, 64, 16, rng, 3e-4)]
    mix = serving.make_mix(24, kinds, rate=2000.0, seed=5)
    for i, r in enumerate(mix):
        r.id = 100 + i
        r.arrival += 5e-4
    srv = serving.Server(kinds)  # config sizes: t = 1, 100 x 10, 64 GB budget
    st = srv.run(checked + mix, keep=True)
    say("served at mix sizes", {k: st[k] for k in ("requests", "iterations", "iterations_with_several_requests",
                                                   "p50_ms", "p99_ms")})
    assert st["requests"] == len(checked) + len(mix) and st["refused"] == 0
    assert st["iterations_with_several_requests"] > 0
    p = srv.params
    for r in checked:
        k = kinds[r.kind]
        g = k.space.tessellation.geom()
        pts = r.inputs.cpu().numpy()
        got = r.points_out.cpu().numpy()
        T_dev = r.tables[0].cpu().numpy().reshape(k.space.nC, -1)
        sA_dev = r.tables[1].cpu().numpy().reshape(k.space.nC, -1)
        same = orc.calc_T(g, pts, T_dev, sA_dev, p["dt"], p["nTimeSteps"], p["nStepsODEsolver"])
        nbad = int((same != got).any(1).sum())
        Av = srv.avees(r.kind, torch.from_numpy(r.theta).cuda().reshape(1, -1))[0].cpu().numpy()
        sA, T = orc.prep(k.dim, Av.reshape(k.space.nC, -1), k.sign, p["dt"])
        want, _, md = orc.calc_T(g, pts, T, sA, p["dt"], p["nTimeSteps"], p["nStepsODEsolver"], probe=True)
        diff = np.abs(got - want).max(1)
        near = md < 1e-12
        say(r.kind, r.n, "points; given the tables: differing", nbad, "; end to end: max diff", float(diff.max()),
            "near-facet", int(near.sum()), "beyond 1e-10 elsewhere", int((diff[~near] > 1e-10).sum()))
        assert nbad == 0, (r.kind, r.n)
        assert diff[~near].max() <= 1e-10, (r.kind, r.n)
        if r.kind == "3d":
            vol = srv.volumes[r.seed % len(srv.volumes)].cpu().numpy()
            wv = orc.remap3d(got, vol)
            assert np.array_equal(wv.reshape(-1), r.result.cpu().numpy().reshape(-1))
    for r in checked + mix:
        alone = srv.alone(r)
        torch.cuda.synchronize()
        assert torch.equal(alone, r.result), (r.id, r.kind)


def test_oversized_request_is_refused(kinds):
    """A request larger than the whole budget is refused on arrival (it
    could never be admitted); the others are served and the loop ends."""
    from cpabdiffeo_amd import serving
    rng = np.random.default_rng(3)
    small = _request(kinds, 0, "1d", 1000, 0, 1, rng)
    big = _request(kinds, 1, "2d4", 512 * 512, 512, 2, rng)
    srv = serving.Server(kinds, budget_bytes=1 << 20, n_volumes=1,
                         params=dict(dt=0.02, nTimeSteps=20, nStepsODEsolver=10))
    assert srv.bytes_of(big) > srv.budget >= srv.bytes_of(small)
    st = srv.run([small, big], keep=True)
    assert st["requests"] == 1 and st["refused"] == 1
    assert small.result is not None and small.error is None
    assert isinstance(big.error, ValueError) and big.result is None
    st = srv.run([big])
    assert st["requests"] == 0 and st["refused"] == 1 and st["p50_ms"] is None
