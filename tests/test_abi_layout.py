"""The Python mirrors of the C-ABI structs (ctypes Structures, the numpy
record of cpab_work_item) have the field offsets and sizes the header gives
them, as compiled by the host C compiler."""
import ctypes
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

STRUCTS = {
    "cpab_reg2d_desc": None,
    "cpab_batch_inst": None,
    "cpab_batch_sched": None,
    "cpab_work_item": None,
}


def _mirrors():
    from cpabdiffeo_amd.batch import BatchInst, BatchSched
    from cpabdiffeo_amd.registration import RegDesc
    from cpabdiffeo_amd.serving import ITEM

    def ct(cls):
        return {f: getattr(cls, f).offset for f, _ in cls._fields_}, ctypes.sizeof(cls)

    return {
        "cpab_reg2d_desc": ct(RegDesc),
        "cpab_batch_inst": ct(BatchInst),
        "cpab_batch_sched": ct(BatchSched),
        "cpab_work_item": ({n: ITEM.fields[n][1] for n in ITEM.names}, ITEM.itemsize),
    }


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs a host C compiler")
def test_struct_layouts_match_the_header():
    mirrors = _mirrors()
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "cpab_hip.h"', "int main(void) {"]
    for name, (fields, _) in mirrors.items():
        lines.append('printf("%s size %%zu\\n", sizeof(%s));' % (name, name))
        for f in fields:
            lines.append('printf("%s %s %%zu\\n", offsetof(%s, %s));' % (name, f, name, f))
    lines.append("return 0; }")
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "layout.c"), os.path.join(d, "layout")
        open(src, "w").write("\n".join(lines))
        subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), src, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for ln in out:
        if not ln.strip():
            continue
        a = ln.split()
        got[(a[0], a[1])] = int(a[2])
    for name, (fields, size) in mirrors.items():
        assert got[(name, "size")] == size, name
        for f, off in fields.items():
            assert got[(name, f)] == off, (name, f)


def test_work_item_records_keep_their_layout_when_filled_in_slices():
    """The serving loop fills one preallocated record array by slices
    (np.concatenate of records drops the padding and is not used)."""
    from cpabdiffeo_amd.serving import ITEM
    arr = np.zeros(6, dtype=ITEM)
    arr[2:5]["nC0"] = 7
    arr[2:5]["inc_x"] = 0.5
    assert arr.dtype.itemsize == 88 and arr.flags["C_CONTIGUOUS"]
    raw = arr.view(np.uint8).reshape(6, 88)
    assert raw[3, 48:52].view(np.int32)[0] == 7
    assert raw[3, 64:72].view(np.float64)[0] == 0.5


def test_serving_mix_is_reproducible_and_has_the_configured_shares():
    from cpabdiffeo_amd import serving
    kinds = serving.default_kinds()
    a = serving.make_mix(4000, kinds, seed=5)
    b = serving.make_mix(4000, kinds, seed=5)
    assert [(r.kind, r.n, r.arrival) for r in a] == [(r.kind, r.n, r.arrival) for r in b]
    share = {k: sum(r.kind == k for r in a) / len(a) for k in ("1d", "2d4", "2d8", "3d")}
    assert abs(share["1d"] - 0.4) < 0.03 and abs(share["2d4"] + share["2d8"] - 0.4) < 0.03
    assert abs(share["3d"] - 0.2) < 0.03
    rate = len(a) / a[-1].arrival
    assert abs(rate - 2000) < 150
    assert all(1e4 <= r.n <= 1e6 for r in a if r.kind == "1d")
    assert all(r.n == r.side ** 2 and 64 <= r.side <= 1024 for r in a if r.kind.startswith("2d"))
    assert kinds["1d"].geo[0] == 50 and not float(kinds["1d"].geo[3]).is_integer()


def test_heterogeneous_instances_follow_the_workload():
    from cpabdiffeo_amd import batch
    spaces, insts = batch.heterogeneous_instances(nInst=300, seed=2, grids=(4, 8))
    assert set(g for g, _, _ in insts) == {4, 8}
    assert all(1e3 <= n <= 1e6 for _, _, n in insts)
    for g, Av, _ in insts[:5]:
        sp = spaces[g][0]
        assert Av.shape == (sp.nC * 6,)
        # Avees lie in the constraint null space (continuity, zero normal boundary velocity)
        assert np.abs(sp.constraintMat @ Av).max() < 1e-10 * max(1.0, np.abs(Av).max())
        assert np.allclose(sp.B @ sp.Avees2theta(Av), Av, atol=1e-12)
