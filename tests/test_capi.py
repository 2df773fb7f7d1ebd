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
                      (base + "GLOBAL_Ireq= 1; END= 0;", "GLOBAL_Ireq"),
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
