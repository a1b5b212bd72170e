"""Benchmark: device-resident MCMC image registration (one chain, 256 x 256).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Workload: 2D registration by Metropolis on a 256 x 256 FP32 blob image
(40 Gaussian blobs, seed 0); the target is the source warped by a draw from
the squared-exponential CPA prior (4 x 4 rectangles x 4 triangles, nC = 64,
zero normal boundary velocity, length scale 0.1, variance 1) plus
N(0, 0.02^2) noise; Gaussian SSD likelihood with sigma = 0.02; random-walk
proposals N(0, 0.01^2 Sigma_CPA).  One Metropolis step = one proposal:
theta' -> Avees -> expm tables -> calc_T_inv of the 65 536 pixel points
(t = 1, 100 steps x 10 RK2 substeps on facet crossings, FP64) -> bilinear
remap (FP32) -> SSD -> accept, on the GPU, launched from captured hipGraphs.

One bench step is one whole chain as configured: theta = 0, step counter
and acceptance count reset, the log-likelihood of theta = 0 evaluated, then
2 000 Metropolis steps.  W untimed chains run first, then K timed chains;
every timed chain is the same chain (checked bit for bit).  value =
2 000 / (seconds per chain).

Each launch evaluates the next S proposals on the all-accepted path and
decides them in order (speculative prefetch, S the speculation depth): the
chain is the same bit for bit as one step per launch, which is timed too and
reported beside it.  Before anything is timed every candidate depth runs a
whole chain (its graph is captured then and kept), then each candidate is
timed on a whole chain and the fastest is used.

Prints one JSON line: proposals/s, the live roofline of the dominant kernel
(algorithmic FP64 FLOPs of the decided proposals over the launches'
HIP-event durations inside the timed region; the discarded speculative work
is reported beside it), and the CPU oracle timed on this host's cores on
proposals of the same chain.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "CPAB points integrated/sec + MCMC proposals/sec; serving p50/p99 latency"
# Algorithmic FLOPs, counted from the reference source (cpab/gpu/calc_2Dto2D.cu):
# add/sub/mul/div/fmod = 1; compares, min/max, round = 0.
F_LOOKUP = 15            # compute_cell_idx, type I, inside (:55-56, 62-65, 75-76, 104, 124, 132/140)
F_AFFINE = 8             # A_times_b_affine (:27-28)
F_FAST = 2 * F_LOOKUP + F_AFFINE           # 38: one accepted Trel step (:389-391)
F_SUB = F_LOOKUP + 2 * F_AFFINE + 6 + 4    # 41: one RK2 substep (:179-194)
F_REMAP_SSD = 16         # bilinear sample (13) + squared error accumulation (3), per pixel
# one substep of the adjoint sweep as written in cpab_adjoint.h: lookup 15, affine 8,
# midpoint 6, g2 2, L^T g2 6, g1 4, lambda update 10, dJ/dA accumulation 20
F_ADJ_SUB = 71
# one accepted Trel step of the adjoint sweep: dJ/dTrel accumulation 10, E^T lambda 6
F_ADJ_TREL = 16
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector peak
# HBM bytes of one step launch at each speculation depth, from the FETCH_SIZE /
# WRITE_SIZE counter runs scaled by a plain copy (committed summary).
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "k_mcmc_step_registration256_spec.md")
CHAIN_STEPS = 2000       # Metropolis steps of one registration chain
DEPTHS = (16, 24, 32, 40, 48, 56, 64)  # speculation depths tuned over


def measured_traffic(depth):
    """Measured bytes per launch at this depth, or None when not measured."""
    import re
    if not os.path.exists(TRAFFIC_SUMMARY):
        return None
    table = {int(k): int(v) for k, v in
             re.findall(r"traffic_bytes_per_launch\[(\d+)\] = (\d+)", open(TRAFFIC_SUMMARY).read())}
    return table.get(int(depth))


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cpu():
    """(logical CPUs this process may use, CPU model name) of this host."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return n, model


def median_rate(fn, units, reps=5):
    """One warm-up call, then the median over `reps` timed calls of fn():
    (units per second, median seconds)."""
    import numpy as np
    fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    med = float(np.median(ts))
    return units / med, med


def cpu_baseline(problem, thetas, threads, n_single=4):
    """The CPU oracle on this host, on proposals the GPU chain evaluated:
    per proposal prep (sign, dt, expm) + calc_T_inv of the 65 536 pixel points
    + bilinear remap + SSD.  Returns the rates on `threads` threads and on one
    (the first n_single proposals), and the oracle's log-likelihoods."""
    from oracle import oracle
    oracle.lib()
    sp = problem.space
    g = sp.tessellation.geom()
    grid = sp.tessellation.dense_grid_img()
    p = problem.params
    lls = []

    def run(ths, nthreads, keep=False):
        for th in ths:
            Av = problem.B @ th
            sA, T = oracle.prep(2, Av.reshape(sp.nC, -1), -1, p["dt"])
            inv = oracle.calc_T(g, grid, T, sA, p["dt"], p["nTimeSteps"], p["nStepsODEsolver"], nthreads=nthreads)
            w = oracle.remap2d(inv, problem.src, nthreads=nthreads)
            ll = oracle.ssd_ll(w, problem.dst, problem.sigma)
            if keep:
                lls.append(ll)

    run(thetas, threads, keep=True)  # the oracle's log-likelihoods (and the warm-up)
    rate, med = median_rate(lambda: run(thetas, threads), len(thetas))
    rate1, med1 = median_rate(lambda: run(thetas[:n_single], 1), n_single)
    return rate, med, rate1, med1, lls


def cpu_points_per_sec(threads):
    """Config-1 unit of work on the CPU oracle: prep + calc_T_fwd of the
    512 x 512 grid (262 144 points), on `threads` threads and on one."""
    from cpabdiffeo_amd import configs
    from oracle import oracle
    cfg = configs.config1()
    sp = cfg["space"]
    g = sp.tessellation.geom()
    As = cfg["Avees"].reshape(sp.nC, -1)

    def once(n):
        return lambda: oracle.transform(g, sp.nC, As, 1, cfg["pts"], cfg["dt"], cfg["nTimeSteps"], cfg["nStepsODE"],
                                        nthreads=n)

    npts = len(cfg["pts"])
    rate, med = median_rate(once(threads), npts)
    rate1, med1 = median_rate(once(1), npts)
    return {"points_per_sec": rate, "threads": threads, "median_s": med,
            "single_thread_points_per_sec": rate1, "single_thread_median_s": med1}


def config1_extra(torch, ctx):
    """calc_T_fwd of the 512 x 512 grid (262 144 points, one call incl. prep)."""
    from cpabdiffeo_amd import configs
    cfg = configs.config1()
    sp = cfg["space"]
    sp.update_pat(cfg["Avees"])
    pts = torch.from_numpy(cfg["pts"]).cuda()
    out = torch.zeros_like(pts)
    nfb = torch.zeros(len(pts), dtype=torch.int32, device="cuda")
    kw = dict(dt=cfg["dt"], nTimeSteps=100, nStepsODEsolver=10)
    for _ in range(5):
        sp._calc_T(pts=pts, mysign=1, out=out, n_fallback=nfb, **kw)
    torch.cuda.synchronize()
    n = 50
    t0 = time.perf_counter()
    for _ in range(n):
        sp._calc_T(pts=pts, mysign=1, out=out, **kw)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / n
    ctx.prof_enable(True)
    ctx.prof_read()
    for _ in range(n):
        sp._calc_T(pts=pts, mysign=1, out=out, n_fallback=nfb, **kw)
    cnt, ms = ctx.prof_read()
    ctx.prof_enable(False)
    kms = ms / cnt
    fb = int(nfb.sum())
    flops = len(pts) * 100 * F_FAST + fb * 10 * F_SUB
    return {"workload": "calc_T_fwd 512x512 grid, nC=64 tess I, 100 steps x 10 substeps, fp64",
            "points_per_sec": len(pts) / wall, "ms_per_call": wall * 1e3, "kernel_us": kms * 1e3,
            "fallback_steps": fb, "alg_tflops": flops / (kms * 1e-3) / 1e12,
            "frac_fp64_peak": flops / (kms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS}


def config3_extra(torch, ctx):
    """Batching: (a) 1 024 heterogeneous 2D instances (4x4, 8x8 or 16x16 rectangles
    x 4 triangles, own theta, N ~ logU[1e3, 1e6] points U[0,1]^2) in one
    persistent graph-captured launch (+ one prep launch), schedule autotuned,
    against one update_pat + calc_T call per instance; (b) 256 registration
    chains, all proposals of a step in one launch, against the host-driven
    one-proposal-at-a-time loop."""
    import numpy as np
    from cpabdiffeo_amd import batch
    from cpabdiffeo_amd.registration import DeviceMetropolis, HostMetropolis, RegistrationProblem
    spaces, insts = batch.heterogeneous_instances(1024, seed=0)
    plan = batch.BatchPlan(2, "I", [((g, g), (1.0 / g, 1.0 / g), n) for g, _, n in insts])
    As = torch.from_numpy(np.concatenate([Av.reshape(-1, 6) for _, Av, _ in insts])).cuda()
    gen = torch.Generator(device="cuda")
    gen.manual_seed(0)
    pts = torch.rand((plan.nPts, 2), dtype=torch.float64, device="cuda", generator=gen)
    out = torch.empty_like(pts)
    params = dict(dt=0.01, nTimeSteps=100, nStepsODEsolver=10)
    tables = plan.prep(As, 1, params["dt"])
    tune = plan.autotune(tables, pts, out, params["dt"], 100, 10)

    def batched():
        plan.prep(As, 1, params["dt"], tables)
        plan.calc_T(tables, pts, out, params["dt"], 100, 10, use_graph=True)

    batched()
    t_b = min(batch.timed(batched) for _ in range(3))
    ctx.prof_enable(True)
    ctx.prof_read()
    plan.calc_T(tables, pts, out, params["dt"], 100, 10, use_graph=True)
    cnt, ms = ctx.prof_read()
    ctx.prof_enable(False)
    # rejected Trel proposals (each redone by 10 RK2 substeps), counted by the
    # same kernel in a launch of its own: the algorithmic work of the batch
    nfb = torch.zeros(plan.nPts, dtype=torch.int32, device="cuda")
    plan.calc_T(tables, pts, out, params["dt"], 100, 10, use_graph=False, n_fallback=nfb)
    fb = int(nfb.sum(dtype=torch.int64))
    del nfb
    bflops = plan.nPts * 100 * F_FAST + fb * 10 * F_SUB
    out2 = torch.empty_like(pts)
    batch.plain_per_instance(spaces, insts, pts, out2, params)
    t_p = batch.timed(lambda: batch.plain_per_instance(spaces, insts, pts, out2, params))
    same = bool(torch.equal(out, out2))
    a = {"workload": "1024 instances, tess {4,8,16}^2 x 4 triangles, N ~ logU[1e3,1e6] (total %d points), "
                     "calc_T_fwd t=1 100x10 fp64" % plan.nPts,
         "points": plan.nPts, "batched_ms": t_b * 1e3, "batched_points_per_sec": plan.nPts / t_b,
         "batched_kernel_ms": ms / cnt, "fallback_steps": fb, "alg_tflops": bflops / (ms / cnt * 1e-3) / 1e12,
         "frac_fp64_peak": bflops / (ms / cnt * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, "schedule": list(plan.schedule.key()),
         "autotune_ms": {"%d/%d/%d" % k: v for k, v in tune.items()},
         "plain_ms": t_p * 1e3, "plain_points_per_sec": plan.nPts / t_p, "speedup": t_p / t_b,
         "bit_identical_to_plain": same}
    plan.close()
    del pts, out, out2
    torch.cuda.empty_cache()
    prob = RegistrationProblem(n=256, seed=0)
    mh = DeviceMetropolis(prob, nChains=256, scale=0.01, seed=0, spec_depth=1)
    mh.set_theta(np.zeros(mh.d))
    mh.run(16)
    torch.cuda.synchronize()
    n_dev = 48
    t_dev = batch.timed(lambda: mh.run(n_dev))
    mh.close()
    hm = HostMetropolis(prob, scale=0.01, seed=0)
    hm.set_theta(np.zeros(hm.d))
    hm.run(8)
    n_host = 100
    t_host = batch.timed(lambda: hm.run(n_host))
    b = {"workload": "256 chains of the 256x256 registration, one step (256 proposals) per launch",
         "proposals_per_sec": 256 * n_dev / t_dev, "plain_proposals_per_sec": n_host / t_host,
         "speedup": (256 * n_dev / t_dev) / (n_host / t_host)}
    return {"instances_1024": a, "chains_256": b}


def config5_extra(torch, n_req=2000, rate=2000.0):
    """Serving: Poisson arrivals at 2 000 req/s of the 1D / 2D / 3D mix, HBM
    budget 64 GB, continuous batching; then the same request set all at once
    (capacity) and a sample of it one request at a time through the
    single-call path."""
    from cpabdiffeo_amd import serving
    kinds = serving.default_kinds()
    srv = serving.Server(kinds, budget_bytes=64 << 30)
    srv.run(serving.make_mix(100, kinds, rate=rate, seed=99))  # warm the allocator and clocks
    st = srv.run(serving.make_mix(n_req, kinds, rate=rate, seed=0))
    cap = srv.run(serving.make_mix(n_req, kinds, rate=rate, seed=0), realtime=False)
    sample = serving.make_mix(n_req, kinds, rate=rate, seed=0)[:300]
    t0 = time.perf_counter()
    for r in sample:
        srv.alone(r)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    plain_pps = sum(r.n for r in sample) / t_plain
    st.update({"workload": "%d requests, Poisson %g req/s: 40%% 1D (50 cells, 1e4-1e6 pts), 40%% 2D grids 64^2-1024^2 "
                           "(4x4 or 8x8 x4 triangles), 20%% 3D 64^3 volumes (4x4x4 x5 tets, calc_T_inv + trilinear "
                           "remap); t=1 100x10 fp64; HBM budget 64 GB" % (n_req, rate),
               "capacity_points_per_sec": cap["points_per_sec"], "capacity_requests_per_sec": cap["requests_per_sec"],
               "plain_points_per_sec": plain_pps, "capacity_vs_plain": cap["points_per_sec"] / plain_pps})
    return st


def config4_extra(torch, ctx, steps=200, lr=3e-4):
    """Gradient descent on 64 pairs of 128x128 blob images (one theta per pair),
    the loss warping with calc_T_inv (calc_T's map: Trel steps + RK2 redos) and
    the gradient by the fused SSD adjoint of that map; plus a central-difference
    check (h = 1e-6) on every pair whose lookups are the same at theta +- h v."""
    import numpy as np
    from cpabdiffeo_amd import adjoint
    prob = adjoint.PairsProblem(n=128, nInst=64, seed=0)
    adjoint.gradient_descent(prob, 2, lr)  # workspace, warm clocks
    torch.cuda.synchronize()
    ctx.prof_enable(True)
    ctx.prof_read()
    t0 = time.perf_counter()
    theta, hist = adjoint.gradient_descent(prob, steps, lr)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    cnt, ms = ctx.prof_read()
    ctx.prof_enable(False)
    h = hist.cpu().numpy()
    npts = prob.nInst * prob.n * prob.n
    p = prob.params
    nT, nS = p["nTimeSteps"], p["nStepsODEsolver"]
    # redone steps at the first and the last theta of the descent (the count
    # per step moves between them)
    def redone(th):
        As, T = adjoint._tables(torch, th, prob.Bt, -1, prob.space.nC, p["dt"])
        nr = torch.empty(prob.nInst, dtype=torch.float64, device="cuda")
        adjoint.ssd_loss_and_grad_T(prob.space, As, T, prob.src, prob.dst, p, n_redo=nr)
        return float(nr.sum())
    nr = 0.5 * (redone(torch.zeros_like(theta)) + redone(theta))
    # algorithmic work of one gradient: the forward calc_T (accepted Trel steps
    # and RK2 redos) and the adjoint units (16 per accepted step, 71 per redone
    # substep); the backward sweep's replay of the forward is not counted
    flops = (npts * nT * F_FAST + nr * nS * F_SUB) + ((npts * nT - nr) * F_ADJ_TREL + nr * nS * F_ADJ_SUB)
    kus = ms / cnt * 1e3
    # central differences (h = 1e-6, FP64 loss) along one random direction per
    # pair: (1) of the whole loss against the adjoint; (2) over the pixels whose
    # cell lookups and bilinear sample cell are the same at theta + h v and
    # theta - h v (a pixel whose accept/redo decision or substep cell changes
    # jumps by the difference of a Trel step and its RK2 redo: not a
    # derivative), against the adjoint of the loss over the same pixels; the
    # pixels left out are counted
    rng = np.random.default_rng(0)
    th = theta.clone().requires_grad_(True)
    loss = prob.loss(th)
    loss.sum().backward()
    v = torch.from_numpy(rng.standard_normal(theta.shape)).cuda()
    an = (th.grad * v).sum(1)
    with torch.no_grad():
        lp = prob.loss(theta + 1e-6 * v)
        lm = prob.loss(theta - 1e-6 * v)
    fd = (lp - lm) / 2e-6
    rel = ((fd - an).abs() / an.abs().clamp_min(1e-12)).cpu().numpy()
    fd_s, an_s, left = adjoint.fd_check_same_lookups(prob, theta, v)
    rel_s = ((fd_s - an_s).abs() / an_s.abs().clamp_min(1e-12)).cpu().numpy()
    left = left.cpu().numpy()
    # the same gradient bit for bit on a second evaluation
    th2 = theta.clone().requires_grad_(True)
    prob.loss(th2).sum().backward()
    return {"workload": "64 pairs of 128x128 fp32 blob images, nC=64 tess I, one theta per pair, %d gradient-descent "
                        "steps (lr %g), loss warping with calc_T_inv t=1 100x10 fp64 (Trel steps + RK2 redos), "
                        "SSD adjoint of that map" % (steps, lr),
            "ms_per_step": el / steps * 1e3, "steps_per_sec": steps / el, "point_gradients_per_sec": npts * steps / el,
            "adjoint_kernel_us": kus, "adjoint_alg_tflops": flops / (kus * 1e-6) / 1e12,
            "adjoint_frac_fp64_peak": flops / (kus * 1e-6) / 1e12 / FP64_PEAK_TFLOPS,
            "redone_steps_per_gradient": nr,
            "loss_mean_first": float(h[0].mean()), "loss_mean_last": float(h[-1].mean()),
            "fd_rel_err_median_all_pixels": float(np.median(rel)), "fd_rel_err_max_all_pixels": float(rel.max()),
            "fd_pairs_within_1e-4_all_pixels": float((rel < 1e-4).mean()),
            "fd_rel_err_median_same_lookup_pixels": float(np.median(rel_s)),
            "fd_rel_err_max_same_lookup_pixels": float(rel_s.max()),
            "fd_pairs_within_1e-4_same_lookup_pixels": float((rel_s < 1e-4).mean()),
            "fd_pixels_left_out_per_pair_mean": float(left.mean()), "fd_pixels_left_out_per_pair_max": int(left.max()),
            "pixels_per_pair": prob.n * prob.n,
            "gradient_bit_identical_on_rerun": bool(torch.equal(th.grad, th2.grad))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed chains (one bench step = one whole chain)")
    ap.add_argument("--warmup", type=int, default=2, help="untimed chains before the timed ones")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-proposals", type=int, default=32)
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--no-seq", action="store_true", help="skip the one-step-per-launch comparison chain")
    ap.add_argument("--section", choices=["calc_T_fwd", "batching", "gradient_descent", "serving"],
                    help="run only this side workload and print its figures (for per-kernel profiles)")
    ap.add_argument("--gd-steps", type=int, default=200,
                    help="gradient-descent steps of the gradient_descent section (fewer for a counter run)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 256 MiB device copy 5 times (the reference point for the "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        x = torch.empty(64 * 2 ** 20, dtype=torch.float32, device="cuda").fill_(1.0)
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read_per_dispatch": x.numel() * 4,
                          "copy_bytes_written_per_dispatch": x.numel() * 4}), flush=True)
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: the library runs on a single MI355X "
              "(chains and point sets would shard across GPUs without communication)",
              file=sys.stderr)
        sys.exit(2)
    if args.section:
        import torch
        from cpabdiffeo_amd._lib import Context
        ctx = Context.get(0)
        fn = {"calc_T_fwd": lambda: config1_extra(torch, ctx), "batching": lambda: config3_extra(torch, ctx),
              "gradient_descent": lambda: config4_extra(torch, ctx, steps=args.gd_steps), "serving": lambda: config5_extra(torch)}
        print(json.dumps({args.section: fn[args.section]()}), flush=True)
        return

    import numpy as np
    import torch
    from cpabdiffeo_amd._lib import Context
    from cpabdiffeo_amd.registration import DeviceMetropolis, RegistrationProblem

    n = 256
    log("bench: building the 256x256 registration problem")
    problem = RegistrationProblem(n=n, seed=0)
    use_graph = not args.no_graph
    ctx = Context.get(0)
    npts = n * n
    per_prop = npts * (100 * F_FAST + F_REMAP_SSD)  # FLOPs of one proposal's pixels without fallbacks

    def start(mh):
        """theta = 0, step counter and acceptance count reset, ll(theta = 0) evaluated."""
        mh.step_counter.zero_()
        mh.n_accepted.zero_()
        mh.set_theta(np.zeros(mh.d))

    def one_chain(mh):
        start(mh)
        mh.run(CHAIN_STEPS, use_graph=use_graph)
        torch.cuda.synchronize()

    def result(mh):
        return (mh.theta.cpu().numpy()[0].copy(), float(mh.ll.cpu().numpy()[0]), int(mh.n_accepted.cpu()[0]))

    # -- speculation depth: capture and warm every candidate on a whole chain,
    # then time each on a whole chain (best of two) and keep the fastest
    mh = DeviceMetropolis(problem, nChains=1, scale=0.01, seed=0, count_fallbacks=True, spec_depth=0)
    start(mh)
    log("bench: capture and warm-up of each speculation depth %s" % (DEPTHS,))
    for K in DEPTHS:
        mh.set_spec_depth(K)
        one_chain(mh)
    rates = {}
    for K in DEPTHS:
        mh.set_spec_depth(K)
        best = None
        for _ in range(2):
            t0 = time.perf_counter()
            one_chain(mh)
            el = time.perf_counter() - t0
            best = el if best is None else min(best, el)
        rates[K] = CHAIN_STEPS / best
    K_best = max(rates, key=rates.get)
    mh.set_spec_depth(K_best)
    log("bench: depth %d; %d warm-up chains, then %d timed chains of %d steps"
        % (K_best, args.warmup, args.steps, CHAIN_STEPS))
    for _ in range(args.warmup):
        one_chain(mh)

    # -- timed region: args.steps whole chains
    st0 = mh.stats.cpu().numpy()[0].copy()
    nfb0 = int(mh.n_fallback.cpu()[0])
    rep0, dir0 = mh.launch_counts()
    ctx.prof_enable(True)
    ctx.prof_read()
    ends = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        one_chain(mh)
        ends.append(result(mh))
    elapsed = time.perf_counter() - t0
    cnt, ms = ctx.prof_read()
    ctx.prof_enable(False)
    st = mh.stats.cpu().numpy()[0] - st0
    nfb = int(mh.n_fallback.cpu()[0]) - nfb0
    rep1, dir1 = mh.launch_counts()
    mh.close()
    same_every_chain = all(np.array_equal(e[0], ends[0][0]) and e[1] == ends[0][1] and e[2] == ends[0][2]
                           for e in ends)
    chain_s = elapsed / args.steps
    proposals_per_sec = CHAIN_STEPS / chain_s

    # -- the same chain one step per launch; proposals it evaluates, for the CPU leg
    seq = None
    samples = []
    if not args.no_seq:
        log("bench: the same chain one step per launch")
        m1 = DeviceMetropolis(problem, nChains=1, scale=0.01, seed=0, count_fallbacks=True, spec_depth=1)
        one_chain(m1)  # capture, warm
        t0 = time.perf_counter()
        one_chain(m1)
        el1 = time.perf_counter() - t0
        r1 = result(m1)
        seq = {"proposals_per_sec": CHAIN_STEPS / el1, "speedup_of_speculation": el1 / chain_s,
               "same_chain_bit_for_bit": bool(np.array_equal(r1[0], ends[0][0]) and r1[1] == ends[0][1]
                                              and r1[2] == ends[0][2])}
        if not args.no_cpu:
            # every (CHAIN_STEPS / cpu_proposals)-th proposal of the chain and the
            # log-likelihood the GPU found for it
            start(m1)
            done = 0
            stride = max(1, CHAIN_STEPS // args.cpu_proposals)
            for s in range(0, CHAIN_STEPS, stride)[:args.cpu_proposals]:
                if s > done:
                    m1.run(s - done)
                tp, _ = m1.last_proposal()
                m1.run(1)
                _, lp = m1.last_proposal()
                done = s + 1
                samples.append((s, tp.cpu().numpy()[0].copy(), float(lp.cpu().numpy()[0]) * npts))
        m1.close()

    # -- roofline of the dominant kernel (k_mcmc_step + k_mcmc_tables), live over
    # the timed region: HIP events on the launch stream around every graph replay
    # and direct launch.  Algorithmic work = the decided proposals (2 000 steps and
    # the evaluation of theta = 0 per chain) and their rejected-Trel redos; the
    # kernel also reports what it integrated for speculative slots it discarded.
    launches = int(st[2])
    kern_s = ms / cnt * 1e-3
    decided = int(st[3]) + args.steps
    flops_dec = decided * per_prop + nfb * 10 * F_SUB
    flops_int = st[4] * (100 * F_FAST + F_REMAP_SSD) + st[1] * 10 * F_SUB
    achieved = flops_dec / launches / kern_s / 1e12

    extra = gd = batching = serve = None
    if not args.no_extra:
        log("bench: config-1 calc_T_fwd extra")
        extra = config1_extra(torch, ctx)
        log("bench: config-3 batching extra")
        batching = config3_extra(torch, ctx)
        log("bench: config-4 gradient descent extra")
        gd = config4_extra(torch, ctx)
        log("bench: config-5 serving extra")
        serve = config5_extra(torch)

    cpu = None
    if not args.no_cpu and samples:
        threads, model = host_cpu()
        omp = os.environ.get("OMP_NUM_THREADS")
        if omp:
            threads = min(threads, int(omp))
        log("bench: CPU oracle baseline, %d proposals of the chain on %d threads (%s)" % (len(samples), threads, model))
        thetas = [t for _, t, _ in samples]
        rate, med, rate1, med1, lls = cpu_baseline(problem, thetas, threads)
        rel = [abs(a - g) / abs(a) for a, (_, _, g) in zip(lls, samples)]
        c1 = cpu_points_per_sec(threads)
        cpu = {"value": rate, "unit": "proposals/s", "cores": threads, "kind": "port",
               "sample": "%d proposals of the timed chain (every %dth step from 0), each: prep + calc_T_inv of "
                         "65536 pts + remap + SSD; 1 warm-up pass, median of 5 passes (%.2f s) on %d threads; "
                         "single thread: first 4 proposals, %.3f proposals/s (median %.2f s)"
                         % (len(samples), max(1, CHAIN_STEPS // args.cpu_proposals), med, threads, rate1, med1),
               "single_thread_value": rate1, "cpu_model": model, "nproc": host_cpu()[0],
               "build": "gcc -O3 -fopenmp -ffp-contract=off, static schedule over points",
               "ll_max_rel_err_gpu_vs_oracle": max(rel),
               "calc_T_fwd_512x512_points_per_sec": c1}
        if extra is not None:
            extra["cpu_points_per_sec"] = c1["points_per_sec"]
            extra["gpu_vs_cpu"] = extra["points_per_sec"] / c1["points_per_sec"]

    line = {
        "metric": METRIC,
        "value": proposals_per_sec,
        "unit": "proposals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": chain_s * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "mcmc-registration 256x256 fp32 blob image, one chain of %d Metropolis steps from "
                               "theta=0 per bench step, nC=64 tess I, calc_T_inv t=1 100x10 fp64 + bilinear remap + "
                               "SSD sigma=0.02" % CHAIN_STEPS,
                   "chain_steps": CHAIN_STEPS, "points_per_proposal": npts, "spec_depth": K_best,
                   "spec_depth_autotune_proposals_per_sec": {str(k): v for k, v in rates.items()},
                   "graph_replayed_launches": rep1 - rep0, "direct_launches": dir1 - dir0,
                   "accept_ratio": ends[0][2] / float(CHAIN_STEPS),
                   "every_timed_chain_identical": bool(same_every_chain)},
        "points_per_sec": proposals_per_sec * npts,
        "one_step_per_launch": seq,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": measured_traffic(K_best),
                     "kernel": "k_mcmc_step + k_mcmc_tables (S proposals: calc_T_inv + remap + SSD; accept; "
                               "arm next)",
                     "kernel_us": kern_s * 1e6, "launches": launches, "event_launches": int(cnt),
                     "decided_proposals_per_launch": decided / launches,
                     "alg_flops_decided_per_launch": flops_dec / launches,
                     "fallback_steps_decided_per_launch": nfb / launches,
                     "proposals_integrated_per_launch": st[4] / float(npts) / launches,
                     "alg_flops_integrated_per_launch": flops_int / launches,
                     "frac_counting_discarded_slots": flops_int / launches / kern_s / 1e12 / FP64_PEAK_TFLOPS,
                     "end_to_end_frac": flops_dec / elapsed / 1e12 / FP64_PEAK_TFLOPS},
        "cpu_baseline": cpu,
        "calc_T_fwd_512x512": extra,
        "gradient_descent_64x128x128": gd,
        "batching": batching,
        "serving": serve,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
