"""dexcluster_scorebox (dexcluster_scorebox.c:5-227, S_init_dcc S_init.c:51-56,
dcc_scorebox_mxA / _svalue / _srt dcc_scorebox.c:6-130): the dcc half-loop elimination's
out_xdrop order replayed over a grid of (rows kept, columns kept), recording the active
counts, Irem and the mean row and column scores at each grid point.

The reference's schedule differs from lakcluster_scorebox's in three ways, all kept:
  - the inner loop's problem (D_sub) is loaded once and never copied from the outer one's
    (D_ori): it keeps its own drops from one inner loop to the next and never takes the
    outer loop's;
  - the first point runs dexcluster_scorebox_excerpt_1, which computes without
    dcc_lrup_mxdup / dcc_M_mxset, so the drops of its dcc_scorebox_mxA never take effect;
  - every point recomputes the subscores from the masks (no low-rank update).
Pins of the oracle (oracle/lko_oracle.c sb_run with dexcluster_scorebox): the grid's masks
against a Python restatement of that schedule (the current order is the tail of the file's
order, its first rdrop / cdrop entries leave); every grid point's means against the
brute-force half-loop definitions (tests/loopcount.py) on those masks.  The engine
(Run::scorebox_dcc) is held to the oracle bit for bit."""
import os

import numpy as np
import pytest

import loopcount


def _dataset(tmp_path, name, **kw):
    from lakcluster_c_amd import gen
    return gen.make_dataset(str(tmp_path), name, **kw)


def _own_order(tmp_path, oracle, ds):
    text = open(ds.in_path).read().replace("GLOBAL_TEST_TYPE= lakcluster_driver;", "GLOBAL_TEST_TYPE= dexcluster_driver;")
    r = oracle.run_text(text)
    path = str(tmp_path / "xdrop_dcc.txt")
    with open(path, "w") as f:
        for a, b in r.xdrop_a.reshape(-1, 2):
            f.write(f"{int(a)} {int(b)}\n")
    return r, path


def _sb_text(ds, xfile, row_max, row_num, col_max, col_num, out_dir=None):
    text = open(ds.in_path).read().replace("GLOBAL_TEST_TYPE= lakcluster_driver;", "").replace("END= 0;", "").rstrip()
    text += (f"\nGLOBAL_TEST_TYPE= dexcluster_scorebox;\nGLOBAL_scorebox_out_xdrop= {xfile};\n"
             f"GLOBAL_scorebox_row_max= {row_max};\nGLOBAL_scorebox_row_num= {row_num};\n"
             f"GLOBAL_scorebox_col_max= {col_max};\nGLOBAL_scorebox_col_num= {col_num};\n")
    if out_dir is not None:
        text += f"GLOBAL_DIR_NAME= {out_dir};\n"
    return text + "END= 0;\n"


def _schedule(r, grid, oracle):
    """(row set, column set) of every grid point [nr + nc*row_num] under the reference's
    dcc schedule (dexcluster_scorebox_rc / _cr)."""
    row_max, row_num, col_max, col_num = grid
    xa = r.xdrop_a
    rord = [int(a) for a in xa[:, 0] if a >= 0]
    cord = [int(b) for b in xa[:, 1] if b >= 0]
    rows0, cols0 = frozenset(rord), frozenset(cord)
    _, rdrop = oracle.scorebox_xdrop_array(len(rows0), row_max, row_num)
    _, cdrop = oracle.scorebox_xdrop_array(len(cols0), col_max, col_num)

    def drop(state, nr, nc):
        rows, cols = state
        rt, ct = rord[len(rord) - len(rows):], cord[len(cord) - len(cols):]
        return rows - set(rt[:min(nr, len(rows))]), cols - set(ct[:min(nc, len(cols))])

    rows_outer = not (row_max > col_max)
    n_out, n_in = (row_num, col_num) if rows_outer else (col_num, row_num)
    ori = sub = (rows0, cols0)
    out = {}
    for o in range(n_out):
        if o > 0:
            ori = drop(ori, rdrop[o] if rows_outer else 0, 0 if rows_outer else cdrop[o])
        out[o if rows_outer else o * row_num] = ori
        for i in range(1, n_in):
            sub = drop(sub, 0 if rows_outer else rdrop[i], cdrop[i] if rows_outer else 0)
            out[o + i * row_num if rows_outer else i + o * row_num] = sub
    return out


GRIDS = [(59, 4, 299, 3), (40, 3, 120, 4), (59, 3, 200, 1), (30, 1, 250, 3), (50, 3, 20, 3)]


@pytest.mark.parametrize("grid", GRIDS)
def test_oracle_dcc_scorebox_follows_the_schedule(oracle, tmp_path, grid):
    ds = _dataset(tmp_path, "sbd", nA=60, ncols=300, nZ=40, T=3, gamma=0.1, seed=31)
    r, xf = _own_order(tmp_path, oracle, ds)
    sb = oracle.scorebox(_sb_text(ds, xf, *grid))
    row_num, col_num = grid[1], grid[3]
    sched = _schedule(r, grid, oracle)
    d = str(tmp_path / "dir_sbd_in")
    A, Z, Tm, Sm = loopcount.load_problem(d, 40, 3)
    S = oracle.State(open(ds.in_path).read().replace("GLOBAL_TEST_TYPE= lakcluster_driver;",
                                                       "GLOBAL_TEST_TYPE= dexcluster_driver;"))
    tab_ajdk = S.get("ajdk").reshape(15, -1)
    S.close()
    zr = np.arange(Z.shape[0])
    checked = 0
    for nc in range(col_num):
        for nr in range(row_num):
            tab = nr + nc * row_num
            rows, cols = sched[tab]
            assert sb.A_rpop_j_total[nr, nc] == len(rows) and sb.A_cpop_j[nr, nc] == len(cols), (nr, nc)
            assert sb.Irem[nr, nc] == (1 if rows else 0)
            if len(rows) < 2 or len(cols) < 2:
                continue
            ar, kl = np.array(sorted(rows)), np.array(sorted(cols))
            raw = loopcount.halfloop_counts(A, Z, Tm, Sm, tab_ajdk, ar, zr, kl, 3)
            QR, QC = loopcount.final_scores(raw, len(ar), len(zr), len(kl), 3, 0.0, dcc=True)
            den = loopcount.denominators(len(ar), len(zr), len(kl), dcc=True)
            for got, v, pre in ((sb.QR_avg[nr, nc], QR, "QR"), (sb.QC_avg[nr, nc], QC, "QC")):
                scale = max(np.abs(v).max(), sum(np.abs(raw[k]).max() / den[k] for k in raw if k.startswith(pre)))
                assert abs(got - v.mean()) <= 1e-7 * scale, (nr, nc, pre, got, v.mean())
            checked += 1
    assert checked >= 2


def test_oracle_dcc_scorebox_first_point_drops_nothing(oracle, tmp_path):
    """The first grid point computes before its drop takes effect: with row_max below the
    row count it still reports every row."""
    ds = _dataset(tmp_path, "sb0", nA=50, ncols=200, nZ=30, T=1, gamma=0.1, seed=32)
    r, xf = _own_order(tmp_path, oracle, ds)
    sb = oracle.scorebox(_sb_text(ds, xf, 30, 3, 120, 3))
    assert sb.A_rpop_j_total[0, 0] == 50 and sb.A_cpop_j[0, 0] == 200


def test_oracle_dcc_scorebox_short_order_is_an_error(oracle, tmp_path):
    ds = _dataset(tmp_path, "sbs", nA=30, ncols=100, nZ=20, T=1, gamma=0.1, seed=33)
    xf = str(tmp_path / "short.txt")
    with open(xf, "w") as f:
        f.write("3 -1\n-1 5\n")
    with pytest.raises(RuntimeError):
        oracle.scorebox(_sb_text(ds, xf, 20, 2, 50, 2))


ENGINE_GRIDS = [("cc", dict(nA=60, ncols=300, nZ=40, T=3), (59, 4, 299, 3)),
                ("cc_cols_outer", dict(nA=60, ncols=300, nZ=40, T=3), (50, 3, 20, 3)),
                ("caseonly", dict(nA=80, ncols=2500, nZ=0, T=1), (60, 3, 2000, 3)),
                ("masked", dict(nA=70, ncols=400, nZ=50, T=2, row_mask_frac=0.1, col_mask_frac=0.1), (40, 4, 300, 2)),
                ("kappa", dict(nA=60, ncols=300, nZ=40, T=3, kappa_squared=0.4), (30, 3, 100, 4))]


@pytest.mark.gpu
@pytest.mark.parametrize("name,kw,grid", ENGINE_GRIDS)
def test_engine_dcc_scorebox_matches_oracle(oracle, tmp_path, name, kw, grid):
    from lakcluster_c_amd import lkc
    ds = _dataset(tmp_path, name, gamma=0.1, seed=34, **kw)
    _, xf = _own_order(tmp_path, oracle, ds)
    text = _sb_text(ds, xf, *grid)
    print(f"dcc scorebox {name} grid {grid}", flush=True)
    ref = oracle.scorebox(text)
    with lkc.Bcc(text) as run:
        assert run.is_scorebox and run.is_dcc
        sb = run.scorebox()
    np.testing.assert_array_equal(sb.A_rpop_j_total, ref.A_rpop_j_total)
    np.testing.assert_array_equal(sb.A_cpop_j, ref.A_cpop_j)
    np.testing.assert_array_equal(sb.Irem, ref.Irem)
    np.testing.assert_array_equal(sb.QR_avg, ref.QR_avg)
    np.testing.assert_array_equal(sb.QC_avg, ref.QC_avg)


@pytest.mark.gpu
def test_engine_dcc_scorebox_files_and_cli(oracle, tmp_path):
    import subprocess
    from lakcluster_c_amd import lkc
    ds = _dataset(tmp_path, "sbf", nA=60, ncols=300, nZ=40, T=3, gamma=0.1, seed=35)
    _, xf = _own_order(tmp_path, oracle, ds)
    grid = (59, 4, 299, 3)
    ref = oracle.scorebox(_sb_text(ds, xf, *grid))
    a, b = str(tmp_path / "entry"), str(tmp_path / "cli")
    lkc.dexcluster_scorebox(_sb_text(ds, xf, *grid, out_dir=a))
    exe = os.path.join(os.path.dirname(lkc.LIB_PATH), "lkc_driver")
    r = subprocess.run([exe, b], input=_sb_text(ds, xf, *grid).encode(), capture_output=True, timeout=300)
    assert r.returncode == 0, r.stderr.decode()
    names = {"A_rpop_j_total": ref.A_rpop_j_total, "A_cpop_j": ref.A_cpop_j, "Irem": ref.Irem,
             "QR_avg": ref.QR_avg, "QC_avg": ref.QC_avg}
    for d in (a, b):
        pre = os.path.join(d, "scorebox_r59n4_c299n3")
        for nm, exp in names.items():
            np.testing.assert_array_equal(lkc.read_mda(f"{pre}_{nm}.mda"), exp)
