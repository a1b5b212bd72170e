"""Maximum sizes: calc_T, calc_v and calc_cell_idx over more than 2^31
points, through the C ABI.

The point indices of one launch cross the signed 32-bit range (2^31 + 4 099
points, 34 GB per (N, 2) float64 array), and the last wave's range is
partial.  Sampled points -- the first ones, those around index 2^31, the last
ones and random ones -- are compared bit for bit with the oracle started from
the same per-cell tables, the fallback counts included.
"""
import numpy as np
import pytest

from test_gpu_parity import random_field, say, space_for

pytestmark = pytest.mark.gpu

N = (1 << 31) + 4099


def sample_indices(rng):
    m = 1 << 31
    idx = np.concatenate([np.arange(0, 1500), np.arange(m - 1500, m + 1500), np.arange(N - 1500, N),
                          rng.integers(0, N, 1500)])
    return np.unique(idx)


def test_calc_T_v_cells_beyond_2_31_points(gpu, orc):
    torch = gpu
    free, _ = torch.cuda.mem_get_info()
    need = N * (16 + 16 + 4) + (4 << 30)
    if free < need:
        pytest.skip("needs %.0f GB of free device memory, %.0f GB free" % (need / 2**30, free / 2**30))
    say("calc_T / calc_v / cell_idx over", N, "points")
    sp = space_for([0, 0], [1, 1], (4, 4), "I")
    g = sp.tessellation.geom()
    Av = random_field(sp, 31)
    sA, T = orc.prep(2, Av.reshape(sp.nC, -1), -1, 0.01)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(32)
    pos0 = torch.rand((N, 2), dtype=torch.float64, device="cuda", generator=gen)
    out = torch.empty_like(pos0)
    nfb = torch.zeros(N, dtype=torch.int32, device="cuda")
    calcs = sp._gpu_calcs
    tdev = torch.from_numpy(np.ascontiguousarray(T)).cuda()
    adev = torch.from_numpy(np.ascontiguousarray(sA)).cuda()
    calcs.calc_T(None, None, tdev, adev, pos0, 0.01, 100, 10, out, 2, 2, sp.nCs, sp.incs, n_fallback=nfb)
    torch.cuda.synchronize()
    idx = sample_indices(np.random.default_rng(33))
    it = torch.from_numpy(idx).cuda()
    p_s = pos0[it].cpu().numpy()
    want, nfb_w, _ = orc.calc_T(g, p_s, T, sA, 0.01, 100, 10, probe=True)
    got = out[it].cpu().numpy()
    bad = np.nonzero((got != want).any(1))[0]
    assert bad.size == 0, (bad.size, idx[bad[:5]])
    assert np.array_equal(nfb[it].cpu().numpy(), nfb_w)
    assert nfb_w.sum() > 0  # the fallback path ran on the sample
    say("calc_T ok; fallbacks on the sample", int(nfb_w.sum()))

    # velocities into the same output buffer, cell ids into the counts' buffer
    sp.update_pat(Av)
    calcs.calc_velocities(None, None, torch.from_numpy(Av.reshape(sp.nC, -1)).cuda(), pos0, out, 2, 2,
                          sp.nCs, sp.incs)
    calcs.calc_cell_idx(None, None, pos0, nfb, 2, 2, sp.nCs, sp.incs)
    torch.cuda.synchronize()
    assert np.array_equal(out[it].cpu().numpy(), orc.calc_v(g, p_s, Av.reshape(sp.nC, -1)))
    assert np.array_equal(nfb[it].cpu().numpy(), orc.calc_cell_idx(g, p_s))
    say("calc_v and cell_idx ok")
    del pos0, out, nfb
    torch.cuda.empty_cache()
