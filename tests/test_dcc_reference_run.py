"""The reference's committed dcc run, reproduced byte for byte.

`dir_test/out_trace.txt`, `out_xdrop_a.txt` and `out_xdrop_b.txt` of the reference are the
output of `dcc_sumscores_test` (dcc_sumscores.c:700-747) for `dir_in/dcc_sumscores_error.in`:
two bins, Y/W genetic controls with 4800 columns, 3 covariate columns, three iterations,
on the problem `dcc_init` (dcc.c:1217-1275) draws with glibc `rand()` from its initial
state.  The files are kept as fixtures under tests/golden/dcc_sumscores_2bins/ (data the
reference's own test directory holds).  The oracle's multi-bin dcc restatement
(oracle/lko_dcc_bins.c) regenerates the problem from the same draws and must write the
same three files.

Every shared double stage is in that output: An_ajdk_v, the plain +-1 products, the
D-weighted QC products, xcalc and An_ZtSn_ww at B_MLT 32, the half-loop, the bin sums and
Y/W combination of dcc_sumscores_nrm / _cmb, the first-element-pivot quicksort and
get_xdrop; the trace prints 16 digits of each mean, so a changed summation order, a
dropped term or a wrong normaliser shows in the files.
"""
import os

GOLD = os.path.join(os.path.dirname(__file__), "golden", "dcc_sumscores_2bins")

# dir_in/dcc_sumscores_error.in, restated (GLOBAL_TEST_* keys; gamma and kappa_squared at
# their defaults, lakcluster_ver18.c:43,85-86)
PARAMS = dict(nbins=2, A_n_rows=380, A_n_cols=3100, Z_n_rows=379, Y_n_cols=4800, T_n_cols=3, B_MLT=32,
              niter=3, sparse=1, gamma_type=0, mrand=0.90, gamma=0.125, kappa2=0.0)


def test_oracle_reproduces_reference_dcc_run(oracle, tmp_path):
    tr, xa, xb = oracle.dcc_sumscores_test(str(tmp_path), **PARAMS)
    for f in ("out_trace.txt", "out_xdrop_a.txt", "out_xdrop_b.txt"):
        got = open(os.path.join(tmp_path, f), "rb").read()
        want = open(os.path.join(GOLD, f), "rb").read()
        assert got == want, f
    assert tr.shape == (3, 6) and xa.shape == (151, 2)
    # the drop counts follow get_xdrop at every line
    for i in range(tr.shape[0]):
        r, c = oracle.get_xdrop(tr[i, 1], tr[i, 2], 0.125, 0)
        rows = int((tr[i, 1] - tr[i + 1, 1])) if i + 1 < tr.shape[0] else r
        cols = int((tr[i, 2] - tr[i + 1, 2])) if i + 1 < tr.shape[0] else c
        assert (rows, cols) == (r, c)


def test_reference_dcc_run_is_sensitive(oracle, tmp_path):
    """The fixture discriminates: one more iteration, or the draws of another problem
    size, give different files (the comparison above is not vacuous)."""
    want = open(os.path.join(GOLD, "out_trace.txt"), "rb").read()
    tr, _, _ = oracle.dcc_sumscores_test(str(tmp_path), **{**PARAMS, "niter": 2})
    assert open(os.path.join(tmp_path, "out_trace.txt"), "rb").read() == want[: want.index(b"\n2 ") + 1]
    tr2, _, _ = oracle.dcc_sumscores_test(None, **{**PARAMS, "Y_n_cols": 4799})
    assert tr2.shape == (3, 6) and tr2[0, 3] != tr[0, 3] and tr2[0, 4] != tr[0, 4]
