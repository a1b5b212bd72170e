"""cpab_calc_T_items_noalloc (the work-item calc_T with no allocation: its
queue counter and work list live in the library's static device array)
against cpab_calc_T_items, through the C ABI: the same bits on two calls in
a row, the second with a smaller list, and the size limit rejected on the
host."""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


def _requests(torch, rng, kinds, n_req):
    """Requests of the 2D kinds: tables from cpab_prep_trels, grid points."""
    from cpabdiffeo_amd._lib import Context
    ctx = Context.get(0)
    ctx.bind_stream(torch.cuda.current_stream().cuda_stream)
    reqs = []
    for i in range(n_req):
        k = kinds["2d4" if i % 2 == 0 else "2d8"]
        theta = k.factor @ rng.standard_normal(k.d)
        Av = torch.from_numpy(k.space.B @ theta).cuda()
        T = torch.empty((k.space.nC, 6), dtype=torch.float64, device="cuda")
        S = torch.empty_like(T)
        ctx.call("cpab_prep_trels", 2, k.space.nC, Av.data_ptr(), -1, 0.01, S.data_ptr(), T.data_ptr())
        side = int(rng.choice([40, 64, 100]))
        ar = torch.arange(side, dtype=torch.float64, device="cuda") / side
        yy, xx = torch.meshgrid(ar, ar, indexing="ij")
        pts = torch.stack([xx.reshape(-1), yy.reshape(-1)], 1).contiguous()
        reqs.append((k, T, S, pts))
    return reqs


def _items(reqs, outs, tile):
    from cpabdiffeo_amd.serving import ITEM
    rows = []
    for (k, T, S, pts), out in zip(reqs, outs):
        n = pts.shape[0]
        for b in range(0, n, tile):
            rows.append((T.data_ptr(), S.data_ptr(), pts.data_ptr(), out.data_ptr(), b, min(b + tile, n)) + k.geo)
    arr = np.zeros(len(rows), dtype=ITEM)
    for i, r in enumerate(rows):
        arr[i] = r
    return arr


def _run(torch, ctx, name, reqs, tile):
    from cpabdiffeo_amd.serving import Sched
    outs = [torch.full_like(pts, np.nan) for (_, _, _, pts) in reqs]
    arr = _items(reqs, outs, tile)
    sched = Sched(4096, 256, 2)
    ctx.call(name, 2, 1, len(arr), arr.ctypes.data, 0.01, 100, 10, ctypes.byref(sched))
    torch.cuda.synchronize()
    return outs, len(arr)


def test_noalloc_items_equal_items_two_calls(gpu):
    torch = gpu
    from cpabdiffeo_amd import serving
    from cpabdiffeo_amd._lib import Context
    kinds = serving.default_kinds()
    ctx = Context.get(0)
    ctx.bind_stream(torch.cuda.current_stream().cuda_stream)
    rng = np.random.default_rng(0)
    first = _requests(torch, rng, kinds, 12)
    second = _requests(torch, rng, kinds, 3)
    for reqs, tile in ((first, 256), (second, 1000)):
        want, n = _run(torch, ctx, "cpab_calc_T_items", reqs, tile)
        got, n2 = _run(torch, ctx, "cpab_calc_T_items_noalloc", reqs, tile)
        say("items", n, "requests", len(reqs))
        assert n == n2 and n <= 4096
        for w, g in zip(want, got):
            assert not torch.isnan(g).any()
            assert torch.equal(w, g)


def test_noalloc_items_size_limit(gpu):
    torch = gpu
    from cpabdiffeo_amd import serving
    from cpabdiffeo_amd._lib import Context
    kinds = serving.default_kinds()
    ctx = Context.get(0)
    rng = np.random.default_rng(1)
    reqs = _requests(torch, rng, kinds, 1)
    outs = [torch.empty_like(reqs[0][3])]
    arr = _items(reqs, outs, 1)  # one item per point: 1 600 to 10 000 items
    while len(arr) <= 4096:
        arr = np.concatenate([arr, arr]).astype(serving.ITEM)
    with pytest.raises(ValueError):
        ctx.call("cpab_calc_T_items_noalloc", 2, 1, len(arr), arr.ctypes.data, 0.01, 100, 10, None)
