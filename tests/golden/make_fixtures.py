"""Write the small committed fixtures of tests/golden/.

    python tests/golden/make_fixtures.py

* cfg1_small.npz : config-1 field (Avees, seed 0) on 4096 points (every
  64th point of the 512 x 512 grid), with the oracle's Trels, signed As,
  calc_T_fwd / calc_T_inv outputs, per-point fallback counts and facet
  distances.
* burkardt_expm_kat.json is data copied from the reference's known-answer
  output (cpab/cpaNd/cy/libc/test_matrix_exponential_prb_output.txt:11-182);
  this script does not write it.

The inputs come from this project's own generators (cpabdiffeo_amd.configs)
and the expected outputs from the CPU oracle; the fixtures freeze both so a
change in either shows up as a test failure.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from cpabdiffeo_amd import configs  # noqa: E402
from oracle import oracle  # noqa: E402


def main():
    cfg = configs.config1()
    sp = cfg["space"]
    geom = sp.tessellation.geom()
    pts = cfg["pts"][::64].copy()
    As = cfg["Avees"].reshape(sp.nC, -1)
    out = {"Avees": cfg["Avees"], "pts": pts, "dt": cfg["dt"], "nTimeSteps": cfg["nTimeSteps"],
           "nStepsODE": cfg["nStepsODE"], "incs": sp.incs, "nCs": sp.nCs}
    for name, sign in (("fwd", 1), ("inv", -1)):
        sA, T = oracle.prep(2, As, sign, cfg["dt"])
        y, nfb, md = oracle.calc_T(geom, pts, T, sA, cfg["dt"], cfg["nTimeSteps"], cfg["nStepsODE"],
                                   probe=True)
        out.update({"signedAs_" + name: sA, "Trels_" + name: T, "T_" + name: y,
                    "nfb_" + name: nfb, "mindist_" + name: md})
    np.savez_compressed(os.path.join(HERE, "cfg1_small.npz"), **out)
    print("wrote cfg1_small.npz: %d points, fallbacks fwd %d inv %d" % (
        len(pts), out["nfb_fwd"].sum(), out["nfb_inv"].sum()))


if __name__ == "__main__":
    main()
