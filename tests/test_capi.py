"""The C ABI library: it loads, exports every entry point include/lkc.h declares,
and reports errors through status codes (no device is needed for these checks)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "lkc.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lkc_[A-Za-z0-9_]+)\s*\(", text)))


def test_header_declares_the_stage_api():
    names = _declared()
    for n in ("lkc_lakcluster_driver", "lkc_bcc_load", "lkc_bcc_init_lf", "lkc_bcc_iterate", "lkc_bcc_run",
              "lkc_bcc_An_ajdk", "lkc_bcc_lf_ZtSn", "lkc_bcc_lf_YnWt", "lkc_bcc_lf_At_T_YnWt_S_Zn",
              "lkc_bcc_lf_AnZt_S_WnYt", "lkc_bcc_singlestudy_ww", "lkc_bcc_doublestudy_ww", "lkc_bcc_flattenloop",
              "lkc_bcc_sumscores", "lkc_bcc_lrup", "lkc_AnZt_vv", "lkc_An_ajdk_v", "lkc_xcalc", "lkc_An_ZtSn_ww"):
        assert n in names, n


def test_library_exports_every_declared_symbol():
    from lakcluster_c_amd import lkc
    L = lkc.lib()
    missing = [n for n in _declared() if not hasattr(L, n)]
    assert not missing, missing


def test_bsize_matches_reference_padding():
    from lakcluster_c_amd import lkc
    # bsize: round up to 16 bits, then to whole 1920-bit blocks (lakcluster_header.h:262)
    assert lkc.bsize(1) == 240
    assert lkc.bsize(1920) == 240
    assert lkc.bsize(1921) == 480
    assert lkc.bsize(16384) == 2160
    assert lkc.bsize(65536) == 8400


def test_errors_are_status_codes(tmp_path):
    from lakcluster_c_amd import lkc
    with pytest.raises(lkc.LkcError):
        lkc.Bcc("GLOBAL_NBINS= 2; END= 0;")                      # unsupported setting
    with pytest.raises(lkc.LkcError):
        lkc.Bcc("NOT_A_KEY= 1; END= 0;")                          # unknown key
    bad = ("GLOBAL_NBINS= 1; GLOBAL_QR_strategy= YnWt; GLOBAL_QC_strategy= YnWt; GLOBAL_Y_n_cols= 0;"
           f"GLOBAL_A_n_name_= {tmp_path}/missing_n.b16; GLOBAL_A_t_name_= {tmp_path}/missing_t.b16; END= 0;")
    with pytest.raises(lkc.LkcError):
        lkc.Bcc(bad)                                              # unreadable file (or no device)
    L = lkc.lib()
    assert L.lkc_bcc_run(None, -1) != 0
    assert L.lkc_bcc_n_trace(None) == -1


def _load_status(text):
    from lakcluster_c_amd import lkc
    L = lkc.lib()
    h = ctypes.c_void_p()
    rc = L.lkc_bcc_load(text.encode(), ctypes.byref(h))
    if h:
        L.lkc_bcc_free(h)
    return rc, L.lkc_last_error().decode()


def test_config_errors_are_rejected_before_any_device_work():
    """LKC_ERR_CONFIG (-1) for a key updateglobals does not know (the reference exits,
    updateglobals.c:185), for settings this path does not honour, and for the YnWt check of
    the bcc elimination; the dcc elimination reads no strategy (dcc.c:1167)."""
    base = "GLOBAL_NBINS= 1; GLOBAL_QR_strategy= YnWt; GLOBAL_QC_strategy= YnWt; GLOBAL_Y_n_cols= 0; "
    for text, why in ((base + "GLOBAL_gama= 0.1; END= 0;", "unknown key GLOBAL_gama"),
                      ("GLOBAL_NBINS= 1; GLOBAL_QR_strategy= YnWt; GLOBAL_QC_strategy= YnWt store one; END= 0;", "YnWt"),
                      ("GLOBAL_NBINS= 1; END= 0;", "YnWt"),
                      (base.replace("GLOBAL_NBINS= 1", "GLOBAL_NBINS= 2") + "END= 0;", "NBINS")):
        rc, msg = _load_status(text)
        assert rc == -1 and why in msg, (text, rc, msg)
    # keys the reference reads that do not touch this path are accepted: the load gets past
    # the configuration and fails on the missing A file instead
    rc, msg = _load_status(base + "GLOBAL_verbose= 1; GLOBAL_thread_count= 8; GLOBAL_TEST_niter= 3; END= 0;")
    assert rc == -1 and "A_n or A_t" in msg
    # dexcluster_driver with no strategy lines, or with the dcc files' "YnWt store one": past the check
    for qc in ("", "GLOBAL_QC_strategy= YnWt store one; "):
        rc, msg = _load_status("GLOBAL_TEST_TYPE= dexcluster_driver; GLOBAL_NBINS= 1; " + qc + "END= 0;")
        assert rc == -1 and "A_n or A_t" in msg, msg


def test_driver_binary_built():
    from lakcluster_c_amd import lkc
    assert os.access(lkc.DRIVER_PATH, os.X_OK)


def test_get_xdrop_entry_points(oracle):
    """lkc_get_xdrop / _length / _array (xdrop.c:5-111) against the oracle's get_xdrop, and the
    schedule's own invariants: each pass keeps what the previous kept minus its drops, the
    last pass empties the rows or the columns, no earlier pass does."""
    from lakcluster_c_amd import lkc
    L = lkc.lib()
    for gamma, gt in ((0.1, 0), (0.02, 1), (0.05, 2), (0.2, 3), (0.0, 0), (0.01, 0)):
        for r, c in ((64, 400), (2048, 8192), (3, 3), (500, 100), (1, 7), (16384, 65536), (2, 1)):
            assert lkc.get_xdrop(r, c, gamma, gt) == oracle.get_xdrop(r, c, gamma, gt)
            a = lkc.get_xdrop_array(r, c, gamma, gt)
            assert len(a) == L.lkc_get_xdrop_length(r, c, gamma, gt) >= 1
            rr, cc = r, c
            for i, (rd, cd, rk, ck) in enumerate(a):
                assert (rd, cd) == oracle.get_xdrop(rr, cc, gamma, gt)
                rr, cc = rr - rd, cc - cd
                assert (rk, ck) == (rr, cc)
                assert (rr > 0 and cc > 0) == (i + 1 < len(a)), (gamma, gt, r, c, i)
    n = ctypes.c_int(0)
    buf = (ctypes.c_int * 4)()
    assert L.lkc_get_xdrop_array(64.0, 400.0, 0.1, 0, 4, ctypes.byref(n), buf, None, None, None) == -5
    assert n.value == 23
    assert L.lkc_get_xdrop_length(float("nan"), 4.0, 0.1, 0) == -5


def test_stage_entries_are_exported_in_driver_order():
    from lakcluster_c_amd import lkc
    names = _declared()
    for st in lkc.INITIAL_STAGES + lkc.LOOP_STAGES + ("sumscores_dmp", "active"):
        assert "lkc_bcc_" + st in names, st
    # the header lists the loop's stages in the driver's order
    text = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "lkc.h")).read(), flags=re.S)
    pos = [text.index("lkc_bcc_" + st + "(") for st in lkc.LOOP_STAGES]
    assert pos == sorted(pos)


def test_stage_loop_program_built():
    assert os.access(os.path.join(ROOT, "tests", "c", "stage_loop"), os.X_OK)


def _oracle_files(oracle, text, out):
    oracle.run_text(text, out_dir=str(out), write_files=True)
    return {f: open(os.path.join(out, f), "rb").read() for f in ("out_trace.txt", "out_xdrop_a.txt", "out_xdrop_b.txt")}


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["casectrl", "caseonly", "masked", "kappa"])
def test_c_stage_loop_matches_whole_run_and_oracle(datasets, oracle, tmp_path, name):
    """A C program built against include/lkc.h walks the driver's loop one stage entry at a
    time, writing the output files with bcc_sumscores_dmp after each pass; its trace and
    drops equal lkc_bcc_run's byte for byte, and its files equal the oracle's."""
    import subprocess
    ds = datasets[name]
    out = tmp_path / "stages"
    out.mkdir()
    print(f"stage_loop {name}", flush=True)
    p = subprocess.run([os.path.join(ROOT, "tests", "c", "stage_loop"), ds.in_path, str(out)],
                       capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "identical" in p.stdout
    ref = _oracle_files(oracle, open(ds.in_path).read(), tmp_path / "oracle")
    for f, b in ref.items():
        assert open(out / f, "rb").read() == b, f


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["casectrl", "b16_g1", "blocks"])
def test_lkc_driver_files_match_oracle(datasets, oracle, tmp_path, name):
    """`lkc_driver out_dir < file.in`, as `lakcluster_ver18 < file.in`: its three output
    files equal the oracle's byte for byte."""
    import subprocess
    from lakcluster_c_amd import lkc
    ds = datasets[name]
    out = tmp_path / "driver"
    print(f"lkc_driver {name}", flush=True)
    with open(ds.in_path) as fin:
        p = subprocess.run([lkc.DRIVER_PATH, str(out)], stdin=fin, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    ref = _oracle_files(oracle, open(ds.in_path).read(), tmp_path / "oracle")
    for f, b in ref.items():
        assert open(out / f, "rb").read() == b, f


def test_dquicksort_xij_matches_oracle(oracle):
    """lkc_dquicksort_xij, the host sort the elimination falls back to, against the oracle's
    dQuickSort_xij (Quicksort.c:83-141): values, permutation and deepest recursion level
    equal, on lists that take the two-thread path (longer than 16384) with many ties, on
    short ones, and on inputs that reach the 32768-level recursion limit (sorted, reversed
    and constant lists longer than 32769, which the reference leaves partly unsorted)."""
    import numpy as np
    from lakcluster_c_amd import lkc
    rng = np.random.default_rng(3)
    cases = [rng.integers(0, 400, 70000).astype(float),          # ties everywhere, parallel path
             rng.normal(size=70000),                             # distinct, parallel path
             rng.integers(0, 9, 5000).astype(float),             # short, ties
             np.arange(40000, dtype=float),                      # sorted: depth limit reached
             np.arange(40000, 0, -1, dtype=float),               # reversed
             np.full(40000, 2.5),                                # constant
             np.r_[rng.normal(size=30000), np.full(20000, 0.0)],  # a long run of equal scores
             np.array([3.0, 1.0, 3.0, 2.0, 3.0])]
    for v in cases:
        d, ix, depth = lkc.dquicksort_xij(v)
        od, oix = oracle.quicksort(v)
        assert d.tobytes() == od.tobytes()
        np.testing.assert_array_equal(ix, oix)
        assert depth == oracle.quicksort.depth
    assert lkc.dquicksort_xij(np.arange(40000, dtype=float))[2] == 32768


def _with_ireq(text, ireq):
    return text.replace("END= 0;", f"GLOBAL_Ireq= {ireq};\nEND= 0;")


@pytest.mark.parametrize("dcc", [False, True])
def test_oracle_ireq_with_one_bin_is_the_plain_run(datasets, oracle, tmp_path, dcc):
    """GLOBAL_Ireq > 0 sorts each row's scores over the bins and normalises per bin
    (bcc_sumscores.c:358-481, 583-628; dcc_sumscores.c:283-407, 490-541); with one bin that
    sorts a single value and the normalisers are the same, so the files are those of
    GLOBAL_Ireq= 0 byte for byte."""
    text = open(datasets["casectrl"].in_path).read()
    if dcc:
        text = text.replace("GLOBAL_TEST_TYPE= lakcluster_driver;", "GLOBAL_TEST_TYPE= dexcluster_driver;")
    a, b = str(tmp_path / "i0"), str(tmp_path / "i2")
    oracle.run_text(text, out_dir=a, write_files=True)
    oracle.run_text(_with_ireq(text, 2), out_dir=b, write_files=True)
    for f in ("out_trace.txt", "out_xdrop_a.txt", "out_xdrop_b.txt"):
        assert open(f"{a}/{f}").read() == open(f"{b}/{f}").read(), f


@pytest.mark.gpu
def test_engine_ireq_with_one_bin_is_the_plain_run(datasets, oracle, tmp_path):
    from lakcluster_c_amd import lkc
    text = open(datasets["casectrl"].in_path).read()
    dtext = text.replace("GLOBAL_TEST_TYPE= lakcluster_driver;", "GLOBAL_TEST_TYPE= dexcluster_driver;")
    for name, t, run in (("bcc", text, lkc.lakcluster_driver), ("dcc", dtext, lkc.dexcluster_driver)):
        a, b, c = (str(tmp_path / f"{name}_{k}") for k in ("eng1", "ora0", "eng0"))
        run(_with_ireq(t, 1), a)
        run(t, c)
        oracle.run_text(t, out_dir=b, write_files=True)
        for f in ("out_trace.txt", "out_xdrop_a.txt", "out_xdrop_b.txt"):
            assert open(f"{a}/{f}").read() == open(f"{b}/{f}").read() == open(f"{c}/{f}").read(), (name, f)

