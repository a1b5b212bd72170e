"""A device-resident request loop with continuous batching for CPA
transforms (no reference counterpart: the reference transforms one request
per Calcs.calc_T call, cpab/gpu/Calcs.py:454-561).

Requests (1D point transforms, 2D grid warps, 3D volume warps) arrive over
time.  A request is admitted when the in-flight device memory (points,
outputs, tables, volumes) stays within the HBM budget; admission uploads
nothing but its theta: its points are generated on the device, its Avees =
B theta and its tables (signed As, expm Trels) are computed on the device
and stay there.  Its points are cut into tiles.  The loop keeps up to
`inflight` iterations on the GPU; an iteration takes pending tiles of the
admitted requests in arrival order, up to `max_points` points per dimension
class, and runs each class as one persistent launch (cpab_calc_T_items), so
requests that arrive while others run join the next iteration and their work
overlaps on the device.  A request completes when its last tile has run (3D:
then its trilinear remap).  Per point the arithmetic is calc_T's, so a
result is the same alone and in the mix.
"""
import collections
import ctypes
import time

import numpy as np

from ._lib import Context
from .gpuarray import require_gpu
from .space import CpaSpace


# cpab_work_item of include/cpab_hip.h, as a numpy record so an iteration's
# item list is built with array operations
ITEM = np.dtype({"names": ["trels", "As", "pos0", "out", "begin", "end", "nC0", "nC1", "nC2",
                           "inc_x", "inc_y", "inc_z"],
                 "formats": [np.uint64] * 4 + [np.int64] * 2 + [np.int32] * 3 + [np.float64] * 3,
                 "offsets": [0, 8, 16, 24, 32, 40, 48, 52, 56, 64, 72, 80], "itemsize": 88})


class Sched(ctypes.Structure):
    _fields_ = [("tile", ctypes.c_int), ("block", ctypes.c_int), ("chunk", ctypes.c_int)]


VOL = 64  # 3D volumes are VOL^3 float32


class Kind:
    """One request type: its space, prior factor, sign and launch class."""

    def __init__(self, name, space, length_scale, sign, cls):
        self.name, self.space, self.sign, self.cls = name, space, sign, cls
        cov = space.se_prior_cov(length_scale, 1.0)
        w, U = np.linalg.eigh(cov)
        self.factor = U * np.sqrt(np.clip(w, 0.0, None))
        self.d = space.d
        self.dim = space.dim_domain
        self.ne = self.dim * (self.dim + 1)
        T = space.tessellation
        nCs = list(T.nCs) + [0] * (3 - self.dim)
        incs = [float(x) for x in T.incs] + [1.0] * (3 - self.dim)
        self.geo = (int(nCs[0]), int(nCs[1]), int(nCs[2]), incs[0], incs[1], incs[2])


def default_kinds():
    """The served mix's tessellations: 1D 50 intervals on [0, 1] (increment
    from the centre differences, 0.02), 2D 4x4 and 8x8 rectangles x 4
    triangles on [0, 1]^2, 3D 4x4x4 cubes x 5 tetrahedra on [0, 64]^3; zero
    normal velocity on the boundary."""
    return {
        "1d": Kind("1d", CpaSpace([0], [1], [50], zero_v_across_bdry=[True], tess="II", valid_outside=False),
                   0.1, 1, (1, "II")),
        "2d4": Kind("2d4", CpaSpace([0, 0], [1, 1], [4, 4], zero_v_across_bdry=[True] * 2, tess="I",
                                    valid_outside=False), 0.1, 1, (2, "I")),
        "2d8": Kind("2d8", CpaSpace([0, 0], [1, 1], [8, 8], zero_v_across_bdry=[True] * 2, tess="I",
                                    valid_outside=False), 0.1, 1, (2, "I")),
        "3d": Kind("3d", CpaSpace([0] * 3, [VOL] * 3, [4, 4, 4], zero_v_across_bdry=[True] * 3, tess="I",
                                  valid_outside=False), 0.1, -1, (3, "I")),
    }


class Request:
    def __init__(self, rid, kind, arrival, n, side, seed, theta):
        self.id, self.kind, self.arrival, self.n, self.side, self.seed = rid, kind, arrival, n, side, seed
        self.theta = theta
        self.t_admit = self.t_done = None
        self.result = None
        self.error = None  # set when the request is refused (larger than the whole budget)


def make_mix(n_req, kinds, rate=2000.0, seed=0, n1d=(1e4, 1e6), sides=(64, 128, 256, 512, 1024)):
    """Poisson arrivals at `rate`; 40 % 1D (N ~ logU n1d), 40 % 2D grids
    (side from `sides`, tessellation 4x4 or 8x8), 20 % 3D volumes; theta from
    the matching prior."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n_req))
    reqs = []
    for i in range(n_req):
        u = rng.random()
        if u < 0.4:
            k, n, side = "1d", int(round(np.exp(rng.uniform(np.log(n1d[0]), np.log(n1d[1]))))), 0
        elif u < 0.8:
            k = "2d4" if rng.random() < 0.5 else "2d8"
            side = int(rng.choice(sides))
            n = side * side
        else:
            k, n, side = "3d", VOL ** 3, 0
        theta = kinds[k].factor @ rng.standard_normal(kinds[k].d)
        reqs.append(Request(i, k, float(t[i]), n, side, int(rng.integers(1 << 31)), theta))
    return reqs


class Server:
    # Default schedule (measured at the 2 000 req/s mix on MI355X): one point per
    # lane (tile = block = 1 024) keeps a launch's tiles short and spread over
    # every CU; 2^21 points per class per iteration bounds an iteration to
    # about a millisecond.  Larger tiles queue many points behind one block.
    def __init__(self, kinds=None, budget_bytes=64 << 30, max_points=1 << 21, tile=1024, block=1024,
                 inflight=2, params=None, n_volumes=4, seed=0, device=0):
        torch = require_gpu()
        self.torch = torch
        self.kinds = kinds or default_kinds()
        self.budget = int(budget_bytes)
        self.max_points, self.tile, self.inflight = int(max_points), int(tile), int(inflight)
        self.params = params or dict(dt=0.01, nTimeSteps=100, nStepsODEsolver=10)
        self.sched = Sched(tile, block, 1)
        self.ctx = Context.get(device)
        self.B = {k: torch.from_numpy(np.ascontiguousarray(v.space.B)).cuda() for k, v in self.kinds.items()}
        self._grids = {}
        from . import configs
        self.volumes = [torch.from_numpy(configs.blob_image(VOL, 40, seed=seed + i, dims=3).astype(np.float32))
                        .cuda() for i in range(n_volumes)]
        zz, yy, xx = np.mgrid[0:VOL, 0:VOL, 0:VOL]
        self.voxels = torch.from_numpy(np.stack([xx.ravel(), yy.ravel(), zz.ravel()], 1).astype(np.float64)).cuda()

    # ------------------------------------------------------------ request state
    def bytes_of(self, r):
        """Device bytes the request holds while admitted."""
        k = self.kinds[r.kind]
        tables = 2 * k.space.nC * k.ne * 8
        if r.kind == "3d":
            return tables + r.n * 3 * 8 + r.n * 4  # output points + warped volume (voxel grid shared)
        if r.kind == "1d":
            return tables + 2 * r.n * 8
        return tables + r.n * 2 * 8  # output (the side's grid is shared)

    def _grid(self, side):
        g = self._grids.get(side)
        if g is None:
            torch = self.torch
            ar = torch.arange(side, dtype=torch.float64, device="cuda") / side
            yy, xx = torch.meshgrid(ar, ar, indexing="ij")
            g = torch.stack([xx.reshape(-1), yy.reshape(-1)], 1).contiguous()
            self._grids[side] = g
        return g

    def _points(self, r):
        torch = self.torch
        if r.kind == "1d":
            g = torch.Generator(device="cuda")
            g.manual_seed(r.seed)
            return torch.rand((r.n, 1), dtype=torch.float64, device="cuda", generator=g)
        if r.kind == "3d":
            return self.voxels
        return self._grid(r.side)

    def admit(self, reqs):
        """Device state of newly admitted requests, per kind in one go: their
        thetas in one upload, Avees = B theta in one product and the tables of
        all of them in one prep launch; points generated on the device
        (shared read-only grids for 2D and 3D); outputs."""
        torch = self.torch
        by = collections.defaultdict(list)
        for r in reqs:
            by[r.kind].append(r)
        self.ctx.bind_stream(torch.cuda.current_stream().cuda_stream)
        dt = float(self.params["dt"])
        for kind, rs in by.items():
            k = self.kinds[kind]
            th = torch.from_numpy(np.stack([r.theta for r in rs])).cuda()
            m, nce = len(rs), k.space.nC * k.ne
            Av = self.avees(kind, th)
            T = torch.empty((m, nce), dtype=torch.float64, device="cuda")
            S = torch.empty((m, nce), dtype=torch.float64, device="cuda")
            self.ctx.call("cpab_prep_trels", k.dim, m * k.space.nC, Av.data_ptr(), int(k.sign), dt, S.data_ptr(),
                          T.data_ptr())
            for i, r in enumerate(rs):
                r.trels, r.As = T[i], S[i]
                r.pts = self._points(r)
                r.out = torch.empty_like(r.pts)
                r.warped = torch.empty((VOL, VOL, VOL), dtype=torch.float32, device="cuda") if kind == "3d" else None
                r.next = 0
                r.left = -(-r.n // self.tile)

    def finish(self, r):
        """After the last tile: the 3D warp's remap (trilinear, calc_T_inv points)."""
        if r.kind == "3d":
            vol = self.volumes[r.seed % len(self.volumes)]
            self.ctx.call("cpab_remap3d", r.out.data_ptr(), vol.data_ptr(), r.warped.data_ptr(), VOL, VOL, VOL)

    # ------------------------------------------------------------ iterations
    def launch(self, active):
        """One iteration: per launch class, pending tiles in arrival order up
        to max_points, one persistent launch.  Returns [(request, tiles)]."""
        torch = self.torch
        taken = []
        by_cls = collections.OrderedDict()
        for r in active:
            if r.next < r.n:
                by_cls.setdefault(self.kinds[r.kind].cls, []).append(r)
        self.ctx.bind_stream(torch.cuda.current_stream().cuda_stream)
        p = self.params
        tile = self.tile
        for (dim, tess), reqs in by_cls.items():
            plan, budget = [], self.max_points
            for r in reqs:
                cnt = min(-(-(r.n - r.next) // tile), max(1, budget // tile))
                b = r.next + tile * np.arange(cnt, dtype=np.int64)
                e = np.minimum(b + tile, r.n)
                plan.append((r, b, e))
                r.next = int(e[-1])
                budget -= int(e[-1] - b[0])
                taken.append((r, cnt))
                if budget <= 0:
                    break
            # one record array with the C layout (np.concatenate would drop its padding)
            arr = np.empty(sum(len(b) for _, b, _ in plan), dtype=ITEM)
            o = 0
            for r, b, e in plan:
                k = self.kinds[r.kind]
                a = arr[o:o + len(b)]
                a["trels"], a["As"] = r.trels.data_ptr(), r.As.data_ptr()
                a["pos0"], a["out"] = r.pts.data_ptr(), r.out.data_ptr()
                a["begin"], a["end"] = b, e
                a["nC0"], a["nC1"], a["nC2"] = k.geo[0], k.geo[1], k.geo[2]
                a["inc_x"], a["inc_y"], a["inc_z"] = k.geo[3], k.geo[4], k.geo[5]
                o += len(b)
            assert arr.dtype.itemsize == 88 and arr.flags["C_CONTIGUOUS"]
            self.ctx.call("cpab_calc_T_items", dim, 1 if tess == "I" else 2, len(arr), arr.ctypes.data,
                          float(p["dt"]), int(p["nTimeSteps"]), int(p["nStepsODEsolver"]), ctypes.byref(self.sched))
        # requests whose last tile is in this iteration: their remap follows
        # in stream order, before the iteration's completion event
        for r, _ in taken:
            if r.next >= r.n:
                self.finish(r)
        return taken

    def run(self, requests, keep=False, realtime=True):
        """Serve `requests` (sorted by arrival).  Returns a stats dict; with
        keep, every request keeps its input points and result in .inputs and
        .result, and its tables (Trels, signed As) and transformed points in
        .tables and .points_out.  A request that needs more device memory than
        the whole budget can never be admitted: it is refused when it arrives
        (its .error says why, the stats count it) and the others are served."""
        torch = self.torch
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        waiting = collections.deque()
        active = []
        flights = collections.deque()
        used = peak = 0
        iters = shared = 0
        done, refused = [], []
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0
        while pending or waiting or active or flights:
            t = now()
            while pending and (pending[0].arrival <= t or not realtime):
                r = pending.popleft()
                need = self.bytes_of(r)
                if need > self.budget:
                    r.error = ValueError("request %d (%s, %d points) needs %d device bytes, more than the whole "
                                         "budget of %d" % (r.id, r.kind, r.n, need, self.budget))
                    refused.append(r)
                    continue
                waiting.append(r)
            new = []
            while waiting and used + self.bytes_of(waiting[0]) <= self.budget:
                r = waiting.popleft()
                used += self.bytes_of(r)
                new.append(r)
            if new:
                self.admit(new)
                ta = now()
                for r in new:
                    r.t_admit = ta
                active.extend(new)
                peak = max(peak, used)
            can = len(flights) < self.inflight and any(r.next < r.n for r in active)
            if can:
                taken = self.launch(active)
                iters += 1
                shared += len(taken) > 1
                ev = torch.cuda.Event()
                ev.record()
                flights.append((ev, taken))
            progressed = False
            while flights and flights[0][0].query():
                ev, taken = flights.popleft()
                progressed = True
                tdone = now()
                for r, cnt in taken:
                    r.left -= cnt
                    if r.left == 0:
                        r.t_done = tdone
                        used -= self.bytes_of(r)
                        if keep:
                            r.inputs = r.pts
                            r.result = r.warped if r.kind == "3d" else r.out
                            r.tables = (r.trels, r.As)
                            r.points_out = r.out
                        r.pts = r.out = r.trels = r.As = r.warped = None
                        done.append(r)
                active = [r for r in active if r.left > 0]
            if not progressed and not can:
                if flights:
                    time.sleep(2e-5)
                elif pending and realtime:
                    time.sleep(max(0.0, min(pending[0].arrival - now(), 1e-3)))
        makespan = max(now(), 1e-9)
        lat = np.array([r.t_done - r.arrival for r in done])
        pct = lambda a, q: float(np.percentile(a, q) * 1e3) if len(a) else None
        pts = sum(r.n for r in done)
        per_kind = {}
        for k in sorted(set(r.kind for r in done)):
            lk = np.array([r.t_done - r.arrival for r in done if r.kind == k])
            per_kind[k] = {"n": int(len(lk)), "p50_ms": pct(lk, 50), "p99_ms": pct(lk, 99)}
        return {"requests": len(done), "refused": len(refused), "makespan_s": makespan,
                "requests_per_sec": len(done) / makespan, "points_per_sec": pts / makespan, "p50_ms": pct(lat, 50),
                "p99_ms": pct(lat, 99), "peak_bytes": int(peak), "iterations": iters,
                "iterations_with_several_requests": shared, "per_kind": per_kind}

    def autotune(self, requests, candidates=((1024, 1024, 1 << 21), (2048, 1024, 1 << 21), (4096, 1024, 1 << 21),
                                             (4096, 1024, 1 << 22)), budget_p99=None):
        """Serve the same sample of requests (in real time) under each
        candidate schedule (tile, block, max_points) and keep the one with the
        lowest p99 latency; with `budget_p99` (seconds), the highest capacity
        among those within it.  Returns {schedule: stats}."""
        import copy
        res = {}
        for tile, block, mp in candidates:
            self.tile, self.sched, self.max_points = int(tile), Sched(int(tile), int(block), 1), int(mp)
            st = self.run([copy.copy(r) for r in requests])
            cap = self.run([copy.copy(r) for r in requests], realtime=False)
            st["capacity_points_per_sec"] = cap["points_per_sec"]
            res[(tile, block, mp)] = st
        ok = [k for k, v in res.items() if budget_p99 is None or v["p99_ms"] * 1e-3 <= budget_p99]
        if budget_p99 is None or not ok:
            best = min(res, key=lambda k: res[k]["p99_ms"])
        else:
            best = max(ok, key=lambda k: res[k]["capacity_points_per_sec"])
        self.tile, self.sched, self.max_points = best[0], Sched(best[0], best[1], 1), best[2]
        return res

    def avees(self, kind, theta):
        """Avees = B theta on the device for a (m, d) stack of thetas."""
        torch = self.torch
        k = self.kinds[kind]
        D = k.space.nC * k.ne
        Av = torch.empty((theta.shape[0], D), dtype=torch.float64, device="cuda")
        self.ctx.bind_stream(torch.cuda.current_stream().cuda_stream)
        self.ctx.call("cpab_theta2avees", D, k.d, int(theta.shape[0]), self.B[kind].data_ptr(),
                      theta.contiguous().data_ptr(), Av.data_ptr())
        return Av

    def alone(self, r):
        """The same request through the single-call path: update_pat +
        calc_T (+ remap3d), for comparison with its result in the mix."""
        torch = self.torch
        k = self.kinds[r.kind]
        sp = k.space
        pts = r.inputs if getattr(r, "inputs", None) is not None else self._points(r)
        out = torch.empty_like(pts)
        sp.update_pat(self.avees(r.kind, torch.from_numpy(r.theta).cuda().reshape(1, -1))[0].cpu().numpy())
        sp._calc_T(pts=pts, mysign=k.sign, out=out, **self.params)
        if r.kind != "3d":
            return out
        warped = torch.empty((VOL, VOL, VOL), dtype=torch.float32, device="cuda")
        vol = self.volumes[r.seed % len(self.volumes)]
        self.ctx.bind_stream(torch.cuda.current_stream().cuda_stream)
        self.ctx.call("cpab_remap3d", out.data_ptr(), vol.data_ptr(), warped.data_ptr(), VOL, VOL, VOL)
        return warped
