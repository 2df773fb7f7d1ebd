"""Writes tests/golden/oracle_caseonly_4096x16384.npz: the CPU oracle's whole elimination of
the 4096 x 16384 case-only problem (A only, p_j ~ U[0.05, 0.5], one planted bicluster
of round(4096^0.55) rows, gamma 0.05, B_MLT 32, T = 1, seed 1): its trace, xdrop lists
and bit-op count.  The generator's output is reproduced from the seed, so only the
results are stored.  The oracle takes about 2-3 minutes on 8 threads.
Run from the repository root:  python tests/golden/make_caseonly_golden.py
"""
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from lakcluster_c_amd import gen  # noqa: E402

CASE_ONLY_4096 = dict(nA=4096, ncols=16384, nZ=0, T=1, gamma=0.05, B=32, seed=1)


def main():
    from oracle import lko
    with tempfile.TemporaryDirectory() as root:
        ds = gen.make_dataset(root, "caseonly_4096x16384", **CASE_ONLY_4096)
        r = lko.run_text(open(ds.in_path).read())
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle_caseonly_4096x16384.npz"),
                        trace=r.trace, xdrop_a=r.xdrop_a.astype(np.int16), xdrop_b=r.xdrop_b.astype(np.int16),
                        ops=np.array([r.ops], dtype=np.int64))


if __name__ == "__main__":
    main()
