"""Many independent calc_T problems in one persistent launch
(cpab_batch_*), with a schedule autotuner.

The reference runs one Calcs.calc_T per instance (cpab/gpu/Calcs.py:454-561)
after one prepare_signedSqAs_and_Tlocals_for_gpu per instance
(cpab/cpaNd/_CpaCalcs.py:227-313); here the instances' tables are prepared by
one launch and all their points integrated by another, per point
bit-identical to the single-instance call.
"""
import ctypes
import time

import numpy as np

from ._lib import Context, check, lib
from .gpuarray import require_gpu
from .tessellation import tess_code


class BatchInst(ctypes.Structure):
    """cpab_batch_inst of include/cpab_hip.h."""
    _fields_ = [("nC0", ctypes.c_int), ("nC1", ctypes.c_int), ("nC2", ctypes.c_int),
                ("inc_x", ctypes.c_double), ("inc_y", ctypes.c_double), ("inc_z", ctypes.c_double),
                ("nPts", ctypes.c_int64)]


class BatchSched(ctypes.Structure):
    """cpab_batch_sched: points per tile, threads per block, tiles per queue grab."""
    _fields_ = [("tile", ctypes.c_int), ("block", ctypes.c_int), ("chunk", ctypes.c_int)]

    def key(self):
        return (self.tile, self.block, self.chunk)


# Candidate schedules the autotuner times: tiles from one point per lane of a
# 1 024-thread block to whole medium instances, blocks of 4 to 16 waves, 1 to
# 8 tiles per queue grab (larger grabs reuse an instance's tables in LDS longer).
CANDIDATES = [(t, b, c) for t in (1024, 4096, 16384, 65536) for b in (256, 1024) for c in (1, 8)]

_TUNED = {}  # signature -> best schedule key, per process


class BatchPlan:
    """A fixed list of instances: geometry of each (one dim and tess for the
    batch) and its point count.  `instances` are (nCs, incs, nPts) triples."""

    def __init__(self, dim, tess, instances, device=0):
        self.dim, self.tess = dim, tess
        arr = (BatchInst * len(instances))()
        for i, (nCs, incs, n) in enumerate(instances):
            nCs = list(nCs) + [0] * (3 - len(nCs))
            incs = list(incs) + [1.0] * (3 - len(incs))
            arr[i] = BatchInst(int(nCs[0]), int(nCs[1]), int(nCs[2]), float(incs[0]), float(incs[1]),
                               float(incs[2]), int(n))
        self._ctx = Context.get(device)
        h = ctypes.c_void_p()
        check(lib().cpab_batch_create(self._ctx.handle, dim, tess_code(tess), len(instances), arr, ctypes.byref(h)),
              self._ctx.handle)
        self._h = h
        nc, npts, nt, mx = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int()
        check(lib().cpab_batch_sizes(h, ctypes.byref(nc), ctypes.byref(npts), ctypes.byref(nt), ctypes.byref(mx)),
              self._ctx.handle)
        self.nInst = len(instances)
        self.nCells, self.nPts, self.tableDoubles, self.maxNC = nc.value, npts.value, nt.value, mx.value
        self.npts_each = np.array([n for _, _, n in instances], dtype=np.int64)
        self.offsets = np.concatenate([[0], np.cumsum(self.npts_each)])
        self.schedule = BatchSched(16384, 1024, 4)

    def _bind(self):
        torch = require_gpu()
        self._ctx.bind_stream(torch.cuda.current_stream().cuda_stream)

    def prep(self, As, sign, dt, tables=None):
        """Signed tables and Trels of every instance from the concatenated
        unsigned As (nCells, dim(dim+1))."""
        torch = require_gpu()
        ne = self.dim * (self.dim + 1)
        if tuple(As.shape) != (self.nCells, ne):
            raise ValueError((tuple(As.shape), (self.nCells, ne)))
        if tables is None:
            tables = torch.empty(self.tableDoubles, dtype=torch.float64, device=As.device)
        self._bind()
        check(lib().cpab_batch_prep(self._h, As.contiguous().data_ptr(), int(sign), float(dt), tables.data_ptr()),
              self._ctx.handle)
        return tables

    def calc_T(self, tables, pos0, out, dt, nTimeSteps, nStepsODE, sched=None, use_graph=True, n_fallback=None):
        if tuple(pos0.shape) != (self.nPts, self.dim) or tuple(out.shape) != tuple(pos0.shape):
            raise ValueError((tuple(pos0.shape), tuple(out.shape)))
        s = sched if sched is not None else self.schedule
        self._bind()
        check(lib().cpab_batch_calc_T(self._h, tables.data_ptr(), pos0.data_ptr(), out.data_ptr(), float(dt),
                                      int(nTimeSteps), int(nStepsODE), ctypes.byref(s), 1 if use_graph else 0,
                                      n_fallback.data_ptr() if n_fallback is not None else None), self._ctx.handle)
        return out

    def signature(self, nTimeSteps, nStepsODE):
        """What the best schedule depends on: dim, tess, instance count, the
        point-count distribution (log2 buckets) and the largest tessellation."""
        hist = np.bincount(np.log2(np.maximum(self.npts_each, 1)).astype(int), minlength=32)[:32]
        return (self.dim, self.tess, self.nInst, self.maxNC, nTimeSteps, nStepsODE, tuple(hist.tolist()))

    def autotune(self, tables, pos0, out, dt, nTimeSteps, nStepsODE, candidates=CANDIDATES, reps=2):
        """Time every candidate schedule on this batch (HIP events around the
        launch), keep the fastest for this plan and for plans with the same
        signature.  Returns {schedule: best ms}."""
        torch = require_gpu()
        sig = self.signature(nTimeSteps, nStepsODE)
        if sig in _TUNED:
            self.schedule = BatchSched(*_TUNED[sig])
            return {}
        times = {}
        for cand in candidates:
            s = BatchSched(*cand)
            self.calc_T(tables, pos0, out, dt, nTimeSteps, nStepsODE, sched=s, use_graph=False)  # warm
            best = float("inf")
            for _ in range(reps):
                self._ctx.prof_enable(True)
                self._ctx.prof_read()
                self.calc_T(tables, pos0, out, dt, nTimeSteps, nStepsODE, sched=s, use_graph=False)
                cnt, ms = self._ctx.prof_read()
                self._ctx.prof_enable(False)
                best = min(best, ms / max(cnt, 1))
            times[cand] = best
        torch.cuda.synchronize()
        win = min(times, key=times.get)
        _TUNED[sig] = win
        self.schedule = BatchSched(*win)
        return times

    def close(self):
        if getattr(self, "_h", None):
            lib().cpab_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def heterogeneous_instances(nInst=1024, seed=0, grids=(4, 8, 16), nmin=1e3, nmax=1e6, length_scale=0.1,
                            variance=1.0):
    """The batching workload: nInst 2D instances on [0,1]^2, each with a
    tessellation drawn uniformly from grids x grids rectangles x 4 triangles,
    its own theta from that tessellation's squared-exponential prior (zero
    normal boundary velocity) and N ~ logU[nmin, nmax] points.  Returns
    (spaces by grid, per-instance (grid, Avees, nPts))."""
    from . import configs
    rng = np.random.default_rng(seed)
    spaces, factors = {}, {}
    for g in grids:
        sp = configs.unit_space((g, g), "I")
        cov = sp.se_prior_cov(length_scale, variance)
        spaces[g] = (sp, cov)
        w, U = np.linalg.eigh(cov)  # theta = U sqrt(w) z, as CpaSpace.sample_theta
        factors[g] = U * np.sqrt(np.clip(w, 0.0, None))
    out = []
    for _ in range(nInst):
        g = int(rng.choice(grids))
        n = int(round(np.exp(rng.uniform(np.log(nmin), np.log(nmax)))))
        sp = spaces[g][0]
        theta = factors[g] @ rng.standard_normal(sp.d)
        out.append((g, sp.B @ theta, n))
    return spaces, out


def plain_per_instance(spaces, insts, pos0, out, params, sign=1):
    """The plain path: per instance, update_pat and one calc_T call (prep +
    integration launch), as the reference's loop over instances would."""
    offs = np.concatenate([[0], np.cumsum([n for _, _, n in insts])])
    for i, (g, Av, n) in enumerate(insts):
        sp = spaces[g][0]
        sp.update_pat(Av)
        sp._calc_T(pts=pos0[offs[i]:offs[i + 1]], mysign=sign, out=out[offs[i]:offs[i + 1]], **params)
    return out


def timed(fn):
    torch = require_gpu()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fn()
    torch.cuda.synchronize()
    return time.perf_counter() - t0
