"""Writes tests/golden/oracle_small.npz: traces and xdrop lists of the CPU oracle on
small seeded problems (the generator's output is reproduced from the seed, so only
the results are stored).  Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from lakcluster_c_amd import gen  # noqa: E402
from oracle import lko  # noqa: E402

CASES = {
    "caseonly_60x300": dict(nA=60, ncols=300, nZ=0, T=1, gamma=0.05, B=32, seed=21),
    "casectrl_50x400_T3": dict(nA=50, ncols=400, nZ=40, T=3, gamma=0.06, B=32, seed=22),
    "masked_45x2100_B24": dict(nA=45, ncols=2100, nZ=35, T=2, gamma=0.08, B=24, seed=23,
                               row_mask_frac=0.15, col_mask_frac=0.1),
}


def run(name, kw, root):
    ds = gen.make_dataset(root, name, **kw)
    r = lko.run_text(open(ds.in_path).read())
    return r


def main():
    out = {}
    with tempfile.TemporaryDirectory() as root:
        for name, kw in CASES.items():
            r = run(name, kw, root)
            out[f"{name}/trace"] = r.trace
            out[f"{name}/xdrop_a"] = r.xdrop_a
            out[f"{name}/xdrop_b"] = r.xdrop_b
            out[f"{name}/ops"] = np.array([r.ops], dtype=np.int64)
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle_small.npz"), **out)


if __name__ == "__main__":
    main()
