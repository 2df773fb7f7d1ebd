"""Writes tests/golden/oracle_casecontrol_2048x8192.npz: the CPU oracle's whole elimination of
the 16384 x 65536 case-control problem's family scaled down to A and Z 2048 x 8192 (same
column model p_j ~ U[0.05, 0.5], a planted bicluster of round(2048^0.55) rows in A only,
T/S 3 covariate columns (intercept + two +-1), gamma 0.02, B_MLT 32, seed 2): its trace,
xdrop lists and bit-op count.  The generator's output is reproduced from the seed, so only
the results are stored.  Run from the repository root:
    python tests/golden/make_casecontrol_golden.py
"""
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from lakcluster_c_amd import gen  # noqa: E402

CASE_CONTROL_2048 = dict(nA=2048, ncols=8192, nZ=2048, T=3, pmin=0.05, pmax=0.5, planted=(0.55,),
                         gamma=0.02, B=32, seed=2)


def main():
    from oracle import lko
    t0 = time.time()
    with tempfile.TemporaryDirectory() as root:
        ds = gen.make_dataset(root, "casecontrol_2048x8192", **CASE_CONTROL_2048)
        r = lko.run_text(open(ds.in_path).read())
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle_casecontrol_2048x8192.npz"),
                        trace=r.trace, xdrop_a=r.xdrop_a.astype(np.int16), xdrop_b=r.xdrop_b.astype(np.int16),
                        ops=np.array([r.ops], dtype=np.int64))
    print(f"{len(r.trace)} trace lines, {len(r.xdrop_a)} xdrop lines, {time.time() - t0:.0f} s")


if __name__ == "__main__":
    main()
