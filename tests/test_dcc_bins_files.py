"""dexcluster_driver (dexcluster_driver.c:5-57) with several bins and Y/W genetic controls,
read from the .in's per-bin file lists (dcc_load, dcc.c:1157-1215).

Pins of the oracle's file-driven multi-bin driver (oracle/lko_dcc_bins.c,
lko_dcc_driver_bins).  Its score stages are the ones the reference's committed
dcc_sumscores run reproduces byte for byte (tests/test_dcc_reference_run.py); what is new
here is the loader (the per-bin lists, the shared column masks, A_p and Y_p from the data,
dcc_load_A_p at dcc.c:875-991) and the run to the end.  Checks:
  - one bin without Y: the same three files, byte for byte, as the one-bin dcc oracle
    (oracle/lko_oracle.c, pinned against the brute-force half-loop definitions in
    tests/test_dcc.py) -- a second, independent restatement of the loader and the loop;
  - Y/W files whose every column is masked off (GLOBAL_Y_n_cind): the same files as no Y/W
    at all, byte for byte (the Y terms vanish exactly, the A terms and A_p are untouched);
  - the same patients split over two or three bins: the half-loop sums over bin pairs are
    bilinear in the bins' row sums, so the scores agree with the one-bin run to the
    B_MLT quantisation (1e-7 relative) and the first step drops the same rows and columns;
  - whole-run invariants: every loaded row and column dropped once, the trace counts
    following get_xdrop, Irem the number of bins with rows left, the b indices the ranks
    among the loaded rows and columns.
The HIP path (lkc_dcc_driver_bins, and lkc_dexcluster_driver / the lkc_driver command line
on the same .in) is held to the oracle's files byte for byte.
"""
import os
import subprocess

import numpy as np
import pytest


def _bins(tmp_path, name, bins, ncols, **kw):
    from lakcluster_c_amd import gen
    return gen.make_bins_dataset(str(tmp_path), name, bins, ncols, **kw)


def _files(d):
    return [open(os.path.join(d, f)).read() for f in ("out_trace.txt", "out_xdrop_a.txt", "out_xdrop_b.txt")]


ONE_BIN = [
    dict(bins=[(90, 0)], ncols=300, T=1, seed=3),
    dict(bins=[(90, 70)], ncols=300, T=1, seed=3),
    dict(bins=[(90, 70)], ncols=300, T=3, seed=3, row_mask_frac=0.1, col_mask_frac=0.1),
    dict(bins=[(80, 0)], ncols=2500, T=2, seed=4, row_mask_frac=0.1),
    dict(bins=[(60, 50)], ncols=400, T=3, seed=5, kappa_squared=0.25, gamma_type=1),
]


@pytest.mark.parametrize("kw", ONE_BIN)
def test_oracle_one_bin_equals_one_bin_oracle(oracle, tmp_path, kw):
    ds = _bins(tmp_path, "one", ny=0, gamma=0.1, **kw)
    text = open(ds.in_path).read()
    a, b = str(tmp_path / "bins"), str(tmp_path / "one")
    oracle.dcc_driver_bins(text, a)
    oracle.run_text(text, out_dir=b, write_files=True)
    assert _files(a) == _files(b)


def test_oracle_masked_y_equals_no_y(oracle, tmp_path):
    from lakcluster_c_amd import gen
    bins = [(40, 30), (50, 40)]
    d0 = _bins(tmp_path, "y0", bins, 300, ny=0, T=3, gamma=0.1, seed=5)
    d1 = _bins(tmp_path, "y1", bins, 300, ny=40, T=3, gamma=0.1, seed=5)
    text = open(d1.in_path).read()
    gen.write_b16(os.path.join(os.path.dirname(d1.in_path), "mc_Y_off.b16"), np.zeros((40, 1), bool))
    a, b, c = str(tmp_path / "o0"), str(tmp_path / "o1"), str(tmp_path / "o2")
    oracle.dcc_driver_bins(open(d0.in_path).read(), a)
    oracle.dcc_driver_bins(text, b)
    oracle.dcc_driver_bins(text.replace("END= 0;", "GLOBAL_Y_n_cind= mc_Y_off.b16;\nEND= 0;"), c)
    assert _files(c) == _files(a)
    # with the Y columns in use the row scores change (the Y/W terms are live)
    assert _files(b)[0] != _files(a)[0]


@pytest.mark.parametrize("ny,T", [(0, 1), (40, 1), (40, 3)])
def test_oracle_bin_split_invariance(oracle, tmp_path, ny, T):
    r = {}
    for tag, bins in (("one", [(90, 70)]), ("two", [(40, 30), (50, 40)]), ("three", [(30, 20), (30, 25), (30, 25)])):
        ds = _bins(tmp_path, tag, bins, 300, ny=ny, T=T, gamma=0.1, seed=5)
        r[tag] = oracle.dcc_driver_bins(open(ds.in_path).read(), str(tmp_path / f"o_{tag}"))
    tr0, xa0, _ = r["one"]
    n0 = int(tr0[0, 1] - tr0[1, 1]) + int(tr0[0, 2] - tr0[1, 2])   # entries dropped by the first step
    assert tr0[0, 5] == 1
    for tag, nbins in (("two", 2), ("three", 3)):
        tr, xa, _ = r[tag]
        # the same schedule of counts; the scores to the quantisation
        np.testing.assert_array_equal(tr[:, 1:3], tr0[:, 1:3])
        np.testing.assert_allclose(tr[:2, 3:5], tr0[:2, 3:5], rtol=1e-7, atol=0)
        # the first step drops the same rows and columns (global row indices over the bins)
        assert sorted(map(tuple, xa[:n0])) == sorted(map(tuple, xa0[:n0]))
        assert tr[0, 5] == nbins


@pytest.mark.parametrize("case", ["y_masks", "mixed_controls", "caseonly_y"])
def test_oracle_bins_run_invariants(oracle, tmp_path, case):
    from lakcluster_c_amd import gen
    kw = dict(y_masks=dict(bins=[(40, 30), (35, 0), (45, 40)], ncols=500, ny=30, T=3, row_mask_frac=0.15,
                           col_mask_frac=0.1, ycol_mask_frac=0.2),
              mixed_controls=dict(bins=[(30, 20), (25, 0), (35, 30)], ncols=500, ny=30, T=3),
              caseonly_y=dict(bins=[(50, 0), (40, 0)], ncols=700, ny=50, T=1))[case]
    ds = _bins(tmp_path, case, gamma=0.1, seed=9, **kw)
    tr, xa, xb = oracle.dcc_driver_bins(open(ds.in_path).read(), str(tmp_path / "o"))
    d = os.path.dirname(ds.in_path)
    nb = len(kw["bins"])
    mr = []
    for b in range(nb):
        f = os.path.join(d, f"mr_A_{b}.b16")
        mr.append(gen.read_b16(f)[:, 0] if os.path.exists(f) else np.ones(kw["bins"][b][0], bool))
    mr = np.concatenate(mr)
    fc = os.path.join(d, "mc_A.b16")
    mc = gen.read_b16(fc)[:, 0] if os.path.exists(fc) else np.ones(kw["ncols"], bool)
    rows = xa[xa[:, 0] >= 0, 0]
    cols = xa[xa[:, 1] >= 0, 1]
    assert sorted(rows) == sorted(np.flatnonzero(mr)) and sorted(cols) == sorted(np.flatnonzero(mc))
    # b indices: the rank of the row (column) among the loaded ones
    rank_r = np.cumsum(mr) - 1
    rank_c = np.cumsum(mc) - 1
    np.testing.assert_array_equal(xb[xb[:, 0] >= 0, 0], rank_r[rows])
    np.testing.assert_array_equal(xb[xb[:, 1] >= 0, 1], rank_c[cols])
    # the trace: iteration numbers, counts before each step, dropped per step by get_xdrop
    assert list(tr[:, 0]) == list(range(len(tr)))
    assert tr[0, 1] == mr.sum() and tr[0, 2] == mc.sum()
    for i in range(len(tr)):
        r, c = oracle.get_xdrop(tr[i, 1], tr[i, 2], 0.1)
        nxt = (tr[i + 1, 1], tr[i + 1, 2]) if i + 1 < len(tr) else (0, 0)
        assert (tr[i, 1] - nxt[0], tr[i, 2] - nxt[1]) == (min(r, tr[i, 1]), min(c, tr[i, 2]))
    # Irem: bins that still hold rows when the step starts
    left = np.concatenate([[0], np.cumsum([int(min(oracle.get_xdrop(t[1], t[2], 0.1)[0], t[1])) for t in tr])])
    order = rows  # global rows in drop order
    bin_of = np.repeat(np.arange(nb), [b[0] for b in kw["bins"]])
    for i in range(len(tr)):
        alive = set(order[left[i]:])
        assert tr[i, 5] == len({bin_of[g] for g in alive})
    assert np.all(np.isfinite(tr[:, 3:5]))


def _write_p(path, p):
    """An A_p / Y_p file as mda_write_d2_r8 writes it: int32 rank 2, the two dimensions, the doubles."""
    with open(path, "wb") as f:
        np.array([2, len(p), 1], dtype=np.int32).tofile(f)
        np.asarray(p, dtype=np.float64).tofile(f)


def test_oracle_bins_p_files(oracle, tmp_path):
    """GLOBAL_A_p_name / GLOBAL_Y_p_name replace the frequencies from the data (dcc_load_A_p,
    dcc.c:936-939, 962-965): files of 0.5 give what GLOBAL_TEST_sparse= 0 gives, byte for byte;
    other values change the run."""
    bins = [(40, 30), (50, 40)]
    ds = _bins(tmp_path, "pf", bins, 2500, ny=2100, T=3, gamma=0.1, seed=8)
    d = os.path.dirname(ds.in_path)
    pa, py = oracle.lib().lko_psize(2500), oracle.lib().lko_psize(2100)
    _write_p(os.path.join(d, "A_p_half.mda"), [0.5] * pa)
    _write_p(os.path.join(d, "Y_p_half.mda"), [0.5] * py)
    _write_p(os.path.join(d, "A_p_var.mda"), np.linspace(0.2, 0.7, pa))
    text = open(ds.in_path).read()
    half = text.replace("END= 0;", "GLOBAL_A_p_name= A_p_half.mda;\nGLOBAL_Y_p_name= Y_p_half.mda;\nEND= 0;")
    dense = text.replace("END= 0;", "GLOBAL_TEST_sparse= 0;\nEND= 0;")
    var = text.replace("END= 0;", "GLOBAL_A_p_name= A_p_var.mda;\nEND= 0;")
    a, b, c, e = (str(tmp_path / n) for n in ("half", "dense", "var", "data"))
    oracle.dcc_driver_bins(half, a)
    oracle.dcc_driver_bins(dense, b)
    oracle.dcc_driver_bins(var, c)
    oracle.dcc_driver_bins(text, e)
    assert _files(a) == _files(b)
    assert _files(c)[0] != _files(e)[0] and _files(c)[0] != _files(a)[0]


def test_bins_config_errors_before_any_device_work(tmp_path):
    """Configuration errors of the multi-bin driver come back as LKC_ERR_CONFIG from the
    loader, which runs on the host before any device call."""
    from lakcluster_c_amd import lkc
    ds = _bins(tmp_path, "err", [(20, 10), (20, 10)], 100, ny=10, T=1, gamma=0.1, seed=2)
    text = open(ds.in_path).read()
    bad = {
        "GLOBAL_Ireq": text.replace("END= 0;", "GLOBAL_Ireq= 1;\nEND= 0;"),
        "bin 1": text.replace("GLOBAL_A_t_name_= A_t_0.b16, A_t_1.b16;", "GLOBAL_A_t_name_= A_t_0.b16;"),
        "different set": text.replace("GLOBAL_Y_t_name_= Y_t_0.b16, Y_t_1.b16;", "GLOBAL_Y_t_name_= Y_t_0.b16;"),
    }
    for why, t in bad.items():
        with pytest.raises(lkc.LkcError) as e:
            lkc.dcc_driver_bins(t, str(tmp_path / "o"))
        assert why in str(e.value), (why, str(e.value))


# ---- the HIP path ----------------------------------------------------------------------

GPU_CASES = {
    "two_y": dict(bins=[(60, 50), (70, 40)], ncols=900, ny=120, T=3, seed=11),
    "three_caseonly_y": dict(bins=[(50, 0), (45, 0), (40, 0)], ncols=1200, ny=64, T=1, seed=12),
    "masks": dict(bins=[(40, 30), (35, 0), (45, 40)], ncols=700, ny=50, T=3, seed=13, row_mask_frac=0.15,
                  col_mask_frac=0.1, ycol_mask_frac=0.2),
    "tmask_kappa": dict(bins=[(50, 40), (40, 50)], ncols=600, ny=40, T=3, seed=14, tcol_mask=[True, False, True],
                        kappa_squared=0.5, gamma_type=1),
    "b16": dict(bins=[(50, 40), (40, 50)], ncols=600, ny=40, T=2, seed=15, B=16),
    "wide": dict(bins=[(120, 100), (130, 90)], ncols=4100, ny=2000, T=3, seed=16),
    "one_bin_y": dict(bins=[(80, 60)], ncols=500, ny=70, T=3, seed=17),
    "p_files": dict(bins=[(50, 40), (60, 30)], ncols=2500, ny=2100, T=2, seed=18),
}


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(GPU_CASES))
def test_engine_dcc_bins_matches_oracle(oracle, tmp_path, name):
    from lakcluster_c_amd import lkc
    kw = dict(GPU_CASES[name])
    ds = _bins(tmp_path, name, gamma=kw.pop("gamma", 0.1), **kw)
    text = open(ds.in_path).read()
    if name == "p_files":
        d = os.path.dirname(ds.in_path)
        _write_p(os.path.join(d, "A_p.mda"), np.linspace(0.15, 0.8, oracle.lib().lko_psize(2500)))
        _write_p(os.path.join(d, "Y_p.mda"), np.linspace(0.6, 0.3, oracle.lib().lko_psize(2100)))
        text = text.replace("END= 0;", "GLOBAL_A_p_name= A_p.mda;\nGLOBAL_Y_p_name= Y_p.mda;\nEND= 0;")
    print(f"dcc bins {name}", flush=True)
    a, b = str(tmp_path / "eng"), str(tmp_path / "ora")
    lkc.dcc_driver_bins(text, a)
    oracle.dcc_driver_bins(text, b)
    assert _files(a) == _files(b)


@pytest.mark.gpu
def test_engine_dcc_bins_one_bin_equals_one_bin_engine(oracle, tmp_path):
    from lakcluster_c_amd import lkc
    ds = _bins(tmp_path, "one", [(90, 70)], 2200, ny=0, T=3, gamma=0.1, seed=3, row_mask_frac=0.1,
               col_mask_frac=0.1)
    text = open(ds.in_path).read()
    a, b, c = str(tmp_path / "bins"), str(tmp_path / "one"), str(tmp_path / "ora")
    lkc.dcc_driver_bins(text, a)
    lkc.dexcluster_driver(text, b)      # one bin, no Y: the one-bin engine
    oracle.run_text(text, out_dir=c, write_files=True)
    assert _files(a) == _files(b) == _files(c)


@pytest.mark.gpu
def test_dexcluster_driver_and_cli_take_the_bins(oracle, tmp_path):
    from lakcluster_c_amd import lkc
    ds = _bins(tmp_path, "route", [(40, 30), (50, 40)], 500, ny=40, T=3, gamma=0.1, seed=21)
    text = open(ds.in_path).read()
    a, b, c = str(tmp_path / "entry"), str(tmp_path / "cli"), str(tmp_path / "ora")
    lkc.dexcluster_driver(text, a)
    exe = os.path.join(os.path.dirname(lkc.LIB_PATH), "lkc_driver")
    r = subprocess.run([exe, b], input=text.encode(), capture_output=True, timeout=300)
    assert r.returncode == 0, r.stderr.decode()
    oracle.dcc_driver_bins(text, c)
    assert _files(a) == _files(b) == _files(c)
