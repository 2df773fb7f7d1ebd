"""lakcluster_scorebox (lakcluster_scorebox.c:5-271, S_init.c:5-89, bcc_scorebox.c:5-133,
282-308): the elimination order of an out_xdrop file is replayed over a grid of (rows
kept, columns kept), recording the active counts, Irem and the mean row and column
scores at each grid point.

Pins of the oracle:
- the drop schedule against the worked example in the reference's docstring
  (bcc_scorebox.c:285-292) and the closed form it implies (xval[n] + 1 indices kept
  at step n once nxols > x_max);
- the counts of every grid point against that schedule (rows follow nr only,
  columns nc only);
- a one-point grid on the elimination's own out_xdrop order against the elimination's
  trace at the same (rows, columns): the same drops, so the same mean scores (the
  elimination recomputing every loop-subscore from scratch after each drop, so that
  both sides compute the same sums);
- every grid point's low-rank-updated scores against recomputation from scratch on the
  same masks (both loop orders: rows outer when row_max <= col_max, columns outer
  otherwise).
The engine is checked against the oracle to the bit."""
import os

import numpy as np
import pytest


def _sb_text(ds, xfile, row_max, row_num, col_max, col_num, out_dir=None):
    text = open(ds.in_path).read().replace("END= 0;", "").rstrip()
    text += (f"\nGLOBAL_TEST_TYPE= lakcluster_scorebox;\nGLOBAL_scorebox_out_xdrop= {xfile};\n"
             f"GLOBAL_scorebox_row_max= {row_max};\nGLOBAL_scorebox_row_num= {row_num};\n"
             f"GLOBAL_scorebox_col_max= {col_max};\nGLOBAL_scorebox_col_num= {col_num};\n")
    if out_dir is not None:
        text += f"GLOBAL_DIR_NAME= {out_dir};\n"
    return text + "END= 0;\n"


def _own_order(tmp_path, oracle, ds, name="xdrop.txt", recompute=False):
    r = oracle.run_text(open(ds.in_path).read(), recompute=recompute)
    path = str(tmp_path / name)
    with open(path, "w") as f:
        for a, b in r.xdrop_a.reshape(-1, 2):
            f.write(f"{int(a)} {int(b)}\n")
    return r, path


def test_xdrop_array_docstring_example(oracle):
    xval, xdrop = oracle.scorebox_xdrop_array(16, 7, 3)
    assert xval.tolist() == [7, 4, 0]
    assert xdrop.tolist() == [8, 3, 4]


@pytest.mark.parametrize("nxols,x_max,x_num", [(16, 7, 3), (150, 99, 5), (900, 511, 7), (40, 40, 4),
                                               (10, 3, 1), (200, 150, 2), (64, 63, 64)])
def test_xdrop_array_closed_form(oracle, nxols, x_max, x_num):
    xval, xdrop = oracle.scorebox_xdrop_array(nxols, x_max, x_num)
    d = x_max / max(1, x_num - 1)
    # round() of the C library: halves away from zero
    exp = [int(np.floor(x_max - n * d + 0.5)) for n in range(x_num)]
    assert xval.tolist() == exp
    assert xdrop[0] == max(0, nxols - (x_max + 1))
    kept = nxols - np.cumsum(xdrop)
    if nxols > x_max:
        np.testing.assert_array_equal(kept, xval + 1)
    assert np.all(xdrop[1:] >= 0)


@pytest.mark.parametrize("name,grid", [("casectrl", (99, 4, 500, 3)), ("casectrl", (120, 3, 60, 4)),
                                       ("caseonly", (150, 3, 300, 3)), ("masked", (80, 3, 700, 2))])
def test_oracle_scorebox_counts(tmp_path, datasets, oracle, name, grid):
    ds = datasets[name]
    r, xf = _own_order(tmp_path, oracle, ds)
    row_max, row_num, col_max, col_num = grid
    rpop0, cpop0 = int(r.trace[0, 1]), int(r.trace[0, 2])
    sb = oracle.scorebox(_sb_text(ds, xf, *grid))
    _, rdrop = oracle.scorebox_xdrop_array(rpop0, row_max, row_num)
    _, cdrop = oracle.scorebox_xdrop_array(cpop0, col_max, col_num)
    er = np.maximum(0, rpop0 - np.cumsum(rdrop))
    ec = np.maximum(0, cpop0 - np.cumsum(cdrop))
    np.testing.assert_array_equal(sb.A_rpop_j_total, np.repeat(er[:, None], col_num, 1))
    np.testing.assert_array_equal(sb.A_cpop_j, np.repeat(ec[None, :], row_num, 0))
    np.testing.assert_array_equal(sb.Irem, (sb.A_rpop_j_total > 0).astype(np.int32))
    assert np.all(np.isfinite(sb.QR_avg)) and np.all(np.isfinite(sb.QC_avg))


@pytest.mark.parametrize("name", ["casectrl", "caseonly", "masked", "b16_g1", "kappa"])
def test_oracle_scorebox_point_matches_trace(tmp_path, datasets, oracle, name):
    ds = datasets[name]
    r, xf = _own_order(tmp_path, oracle, ds, recompute=True)
    tr = r.trace
    picks = sorted({0, 1, len(tr) // 3, (2 * len(tr)) // 3, len(tr) - 2})
    for i in picks:
        rp, cp = int(tr[i, 1]), int(tr[i, 2])
        if rp < 1 or cp < 1:
            continue
        sb = oracle.scorebox(_sb_text(ds, xf, rp - 1, 1, cp - 1, 1))
        assert int(sb.A_rpop_j_total[0, 0]) == rp and int(sb.A_cpop_j[0, 0]) == cp
        assert int(sb.Irem[0, 0]) == int(tr[i, 5])
        # the means differ only in summation order (index order here, score order in the trace)
        np.testing.assert_allclose(sb.QR_avg[0, 0], tr[i, 3], rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(sb.QC_avg[0, 0], tr[i, 4], rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("name,grid", [("casectrl", (99, 4, 500, 3)), ("casectrl", (120, 3, 60, 4)),
                                       ("caseonly", (150, 3, 300, 3)), ("masked", (80, 4, 700, 3)),
                                       ("blocks", (60, 3, 2000, 3))])
def test_oracle_scorebox_update_matches_recompute(tmp_path, datasets, oracle, name, grid):
    ds = datasets[name]
    _, xf = _own_order(tmp_path, oracle, ds)
    text = _sb_text(ds, xf, *grid)
    up, full = oracle.scorebox(text), oracle.scorebox(text, recompute=True)
    np.testing.assert_array_equal(up.A_rpop_j_total, full.A_rpop_j_total)
    np.testing.assert_array_equal(up.A_cpop_j, full.A_cpop_j)
    np.testing.assert_array_equal(up.Irem, full.Irem)
    # the last grid row and column keep a single row or column: there every loop-subscore
    # is a cancellation residue of zero, which the two routes leave at different roundings
    ok = (up.A_rpop_j_total > 2) & (up.A_cpop_j > 2)
    assert ok.sum() >= 4
    np.testing.assert_allclose(up.QR_avg[ok], full.QR_avg[ok], rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(up.QC_avg[ok], full.QC_avg[ok], rtol=1e-6, atol=1e-9)


def test_oracle_scorebox_short_order_is_an_error(tmp_path, datasets, oracle):
    ds = datasets["casectrl"]
    xf = str(tmp_path / "short.txt")
    with open(xf, "w") as f:
        f.write("3 -1\n-1 5\n")
    with pytest.raises(RuntimeError):
        oracle.scorebox(_sb_text(ds, xf, 50, 2, 400, 2))


def test_host_xdrop_array_matches_oracle(oracle):
    from lakcluster_c_amd import lkc
    for args in [(16, 7, 3), (150, 99, 5), (900, 511, 7), (40, 40, 4), (10, 3, 1), (64, 63, 64), (5, 9, 3)]:
        a, b = lkc.scorebox_xdrop_array(*args), oracle.scorebox_xdrop_array(*args)
        np.testing.assert_array_equal(a[0], b[0])
        np.testing.assert_array_equal(a[1], b[1])


# (dataset, grid): rows outer (row_max <= col_max) and columns outer; grids reaching one
# row / one column; a grid wider than the data (all rows dropped at the last point)
ENGINE_CASES = [("casectrl", (99, 4, 500, 3)), ("casectrl", (120, 3, 60, 4)), ("caseonly", (150, 3, 300, 3)),
                ("masked", (80, 4, 700, 3)), ("blocks", (60, 3, 2000, 3)), ("kappa", (90, 3, 200, 5)),
                ("b16_g1", (70, 2, 400, 2)), ("wide_z", (40, 3, 500, 3)), ("tiny", (9, 3, 30, 3))]


@pytest.mark.gpu
@pytest.mark.parametrize("name,grid", ENGINE_CASES)
def test_engine_scorebox_matches_oracle(tmp_path, datasets, oracle, name, grid):
    from lakcluster_c_amd import lkc
    ds = datasets[name]
    _, xf = _own_order(tmp_path, oracle, ds)
    text = _sb_text(ds, xf, *grid)
    print(f"scorebox {name} grid {grid}", flush=True)
    ref = oracle.scorebox(text)
    with lkc.Bcc(text) as run:
        assert run.is_scorebox
        sb = run.scorebox()
    assert sb.dims == grid
    np.testing.assert_array_equal(sb.A_rpop_j_total, ref.A_rpop_j_total)
    np.testing.assert_array_equal(sb.A_cpop_j, ref.A_cpop_j)
    np.testing.assert_array_equal(sb.Irem, ref.Irem)
    np.testing.assert_array_equal(sb.QR_avg, ref.QR_avg)
    np.testing.assert_array_equal(sb.QC_avg, ref.QC_avg)


@pytest.mark.gpu
def test_engine_scorebox_writes_mda_files(tmp_path, datasets, oracle):
    from lakcluster_c_amd import lkc
    ds = datasets["casectrl"]
    _, xf = _own_order(tmp_path, oracle, ds)
    out = str(tmp_path / "sb_out")
    grid = (99, 4, 500, 3)
    text = _sb_text(ds, xf, *grid, out_dir=out)
    ref = oracle.scorebox(text)
    lkc.lakcluster_scorebox(text)
    pre = os.path.join(out, "scorebox_r99n4_c500n3")
    names = {"A_rpop_j_total": (ref.A_rpop_j_total, 4), "A_cpop_j": (ref.A_cpop_j, 4), "Irem": (ref.Irem, 4),
             "QR_avg": (ref.QR_avg, 8), "QC_avg": (ref.QC_avg, 8)}
    for nm, (exp, el) in names.items():
        path = f"{pre}_{nm}.mda"
        assert os.path.getsize(path) == 12 + el * exp.size
        np.testing.assert_array_equal(lkc.read_mda(path), exp)


@pytest.mark.gpu
def test_engine_scorebox_short_order_is_an_error(tmp_path, datasets):
    from lakcluster_c_amd import lkc
    ds = datasets["casectrl"]
    xf = str(tmp_path / "short.txt")
    with open(xf, "w") as f:
        f.write("3 -1\n-1 5\n")
    with lkc.Bcc(_sb_text(ds, xf, 50, 2, 400, 2)) as run:
        with pytest.raises(lkc.LkcError):
            run.scorebox()
