"""Batched calc_T over heterogeneous instances (one persistent launch):
per point bit-identical to the single-instance call and to the oracle given
the same tables, for every schedule, with and without graph replay."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


def _tables_of(plan, tables, i, nC, ne):
    """Trels and signed As of instance i from the batch table buffer."""
    off = 0
    for j in range(i):
        off += 2 * plan._nc[j] * ne
    t = tables[off:off + 2 * nC * ne].reshape(2, nC, ne)
    return t[0], t[1]


@pytest.fixture(scope="module")
def small_batch(gpu):
    import torch
    from cpabdiffeo_amd import batch
    spaces, insts = batch.heterogeneous_instances(nInst=10, seed=3, nmin=50, nmax=3000)
    insts[3] = (insts[3][0], insts[3][1], 0)  # an empty instance
    plan = batch.BatchPlan(2, "I", [((g, g), (1.0 / g, 1.0 / g), n) for g, _, n in insts])
    plan._nc = [spaces[g][0].nC for g, _, _ in insts]
    As = torch.from_numpy(np.concatenate([Av.reshape(-1, 6) for _, Av, _ in insts])).cuda()
    rng = np.random.default_rng(0)
    p = rng.uniform(0, 1, (plan.nPts, 2))
    # a few start points outside the domain and on cell walls / diagonals
    # (multiples of a quarter of every tessellation's increment)
    k = rng.permutation(plan.nPts)[:plan.nPts // 10]
    p[k[:len(k) // 2]] = rng.uniform(-0.2, 1.2, (len(k) // 2, 2))
    p[k[len(k) // 2:]] = rng.integers(0, 65, (len(k) - len(k) // 2, 2)) / 64.0
    pts = torch.from_numpy(p).cuda()
    return dict(spaces=spaces, insts=insts, plan=plan, As=As, pts=pts)


def test_batch_equals_single_calls_and_oracle(small_batch, orc):
    import torch
    from cpabdiffeo_amd import batch
    b = small_batch
    plan, insts, spaces = b["plan"], b["insts"], b["spaces"]
    params = dict(dt=0.01, nTimeSteps=100, nStepsODEsolver=10)
    say("batch of", plan.nInst, "instances,", plan.nPts, "points")
    tables = plan.prep(b["As"], 1, params["dt"])
    out = torch.empty_like(b["pts"])
    nfb = torch.zeros(plan.nPts, dtype=torch.int32, device="cuda")
    plan.calc_T(tables, b["pts"], out, params["dt"], 100, 10, use_graph=False, n_fallback=nfb)
    single = torch.empty_like(b["pts"])
    batch.plain_per_instance(spaces, insts, b["pts"], single, params)
    torch.cuda.synchronize()
    assert torch.equal(out, single)
    tab = tables.cpu().numpy()
    pts = b["pts"].cpu().numpy()
    got, gotf = out.cpu().numpy(), nfb.cpu().numpy()
    for i, (g, _, n) in enumerate(insts):
        if n == 0:
            continue
        sp = spaces[g][0]
        T, sA = _tables_of(plan, tab, i, sp.nC, 6)
        o0, o1 = plan.offsets[i], plan.offsets[i + 1]
        want, wf, _ = orc.calc_T(sp.tessellation.geom(), pts[o0:o1], T, sA, params["dt"], 100, 10, probe=True)
        assert np.array_equal(got[o0:o1], want), i
        assert np.array_equal(gotf[o0:o1], wf), i


@pytest.mark.parametrize("sched", [(64, 64, 1), (1024, 256, 3), (4096, 1024, 1), (100000, 512, 2)])
def test_every_schedule_and_graph_replay_agree(small_batch, sched):
    import torch
    from cpabdiffeo_amd import batch
    b = small_batch
    plan = b["plan"]
    say("schedule", sched)
    tables = plan.prep(b["As"], -1, 0.02)
    ref = torch.empty_like(b["pts"])
    plan.calc_T(tables, b["pts"], ref, 0.02, 50, 10, sched=batch.BatchSched(4096, 256, 4), use_graph=False)
    out = torch.empty_like(b["pts"])
    for rep in range(3):  # graph replays, with eager work between them
        out.fill_(float("nan"))
        plan.calc_T(tables, b["pts"], out, 0.02, 50, 10, sched=batch.BatchSched(*sched), use_graph=True)
        torch.cuda.synchronize()
        assert torch.equal(out, ref), rep


@pytest.mark.parametrize("dim,tess", [(1, "II"), (3, "I"), (2, "II")])
def test_other_dims_match_single_calls(gpu, dim, tess):
    import torch
    from cpabdiffeo_amd import batch
    from cpabdiffeo_amd.space import CpaSpace
    say("batch dim", dim, tess)
    rng = np.random.default_rng(dim)
    # one batch holds power-of-two increments only, or none (1D: 1/50, 1/20)
    grids = {1: [(50,), (20,)], 2: [(2, 4), (4, 4)], 3: [(2, 2, 2), (4, 2, 2)]}[dim]
    spaces = [CpaSpace([0] * dim, [1] * dim, list(g), zero_v_across_bdry=[True] * dim, tess=tess,
                       valid_outside=False) for g in grids]
    insts, As = [], []
    for k in range(5):
        sp = spaces[k % len(spaces)]
        n = int(rng.integers(10, 2000))
        Av = sp.sample_theta(sp.se_prior_cov(0.3, 1.0), rng)
        insts.append((sp, sp.B @ Av, n))
        As.append((sp.B @ Av).reshape(sp.nC, -1))
    plan = batch.BatchPlan(dim, tess, [(sp.nCs, sp.tessellation.incs, n) for sp, _, n in insts])
    pts = torch.from_numpy(rng.uniform(0, 1, (plan.nPts, dim))).cuda()
    tables = plan.prep(torch.from_numpy(np.concatenate(As)).cuda(), 1, 0.02)
    out = torch.empty_like(pts)
    plan.calc_T(tables, pts, out, 0.02, 50, 10, use_graph=True)
    single = torch.empty_like(pts)
    for i, (sp, Av, n) in enumerate(insts):
        sp.update_pat(Av)
        o0, o1 = plan.offsets[i], plan.offsets[i + 1]
        sp._calc_T(pts=pts[o0:o1], mysign=1, out=single[o0:o1], dt=0.02, nTimeSteps=50, nStepsODEsolver=10)
    torch.cuda.synchronize()
    assert torch.equal(out, single)


def test_batch_rejects_mixed_increments(gpu):
    from cpabdiffeo_amd import batch
    with pytest.raises(NotImplementedError):
        batch.BatchPlan(2, "I", [((4, 4), (0.25, 0.25), 10), ((3, 3), (1 / 3, 1 / 3), 10)])


def test_autotune_picks_a_timed_schedule(small_batch):
    import torch
    from cpabdiffeo_amd import batch
    b = small_batch
    plan = b["plan"]
    tables = plan.prep(b["As"], 1, 0.01)
    out = torch.empty_like(b["pts"])
    times = plan.autotune(tables, b["pts"], out, 0.01, 20, 10, candidates=[(1024, 256, 1), (4096, 1024, 4)])
    say("autotune", times)
    assert plan.schedule.key() == min(times, key=times.get)
    assert all(t > 0 for t in times.values())
