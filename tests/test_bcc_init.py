"""bcc_sumscores_test (bcc_sumscores.c:781-880) on the synthetic problem of bcc_init
(bcc.c:1542-1600), one bin without Y/W controls.

bcc_init draws its problem from glibc rand() in a fixed order: bcc_ajdk_init's column masks
(bcc.c:117-174), bcc_single_init_M_An's row masks and matrices (bcc.c:415-520, through
wrap_M_setup_test_excerpt_1 / _0, An_ajdk_v.c:32-98), then bcc_init_A_p's block frequencies
(bcc.c:1315-1330).  With GLOBAL_TEST_mrand > 0 some rows and columns start loaded (_b) but
inactive (_j), which no file-driven run does; the test then runs lakcluster_driver's stages.

Pins of the oracle's restatement (oracle/lko_oracle.c st_init):
  - the whole elimination on the drawn problem against the brute-force loop counts
    (tests/loopcount.py): at every step the trace means equal the means of the normalised,
    corrected and combined brute-force scores over the active rows and columns, and the rows
    and columns dropped are the lowest-scoring ones -- with rows and columns inactive from
    the start;
  - the drawn problem against a second restatement of bcc_init's draw order in numpy on the
    C library's own rand() (ctypes), bit for bit: masks, matrices, covariates, A_p;
  - invariants: every active row and column dropped once, no inactive one, the b indices
    the ranks among the loaded rows and columns.
The HIP path (lkc_bcc_sumscores_test, lkc_bcc_load on the same .in, and the lkc_driver
command line) is held to the oracle's three files byte for byte.
"""
import ctypes
import ctypes.util
import os
import subprocess

import numpy as np
import pytest

import loopcount


def in_text(**kw):
    p = dict(mrand=0.3, A=40, Acols=200, Z=35, Y=0, T=3, niter=0, gamma=0.1, B=32, typ2="error", nbins=1)
    p.update(kw)
    lines = ["GLOBAL_TEST_TYPE= bcc_sumscores;", f"GLOBAL_TEST_TYP2= {p['typ2']};", "GLOBAL_QR_strategy= YnWt;",
             "GLOBAL_QC_strategy= YnWt;", f"GLOBAL_NBINS= {p['nbins']};", f"GLOBAL_TEST_mrand= {p['mrand']};",
             f"GLOBAL_TEST_A_n_rows= {p['A']};", f"GLOBAL_TEST_A_n_cols= {p['Acols']};",
             f"GLOBAL_TEST_Z_n_rows= {p['Z']};", f"GLOBAL_TEST_Y_n_cols= {p['Y']};", f"GLOBAL_TEST_T_n_cols= {p['T']};",
             f"GLOBAL_TEST_niter= {p['niter']};", f"GLOBAL_gamma= {p['gamma']};", f"GLOBAL_B_MLT= {p['B']};"]
    for k in ("kappa_squared", "gamma_type", "TEST_sparse"):
        if k in p:
            lines.append(f"GLOBAL_{k}= {p[k]};")
    return "\n".join(lines + ["END= 0;"]) + "\n"


CASES = {
    "all_active": dict(mrand=0.0),
    "inactive_rows_cols": dict(mrand=0.3),
    "half_inactive_T2": dict(mrand=0.5, T=2),
    "no_controls": dict(mrand=0.2, Z=2, T=1),
    "wide": dict(mrand=0.1, Acols=4000, A=30, Z=25),
    "kappa": dict(mrand=0.2, T=3, kappa_squared=0.4),
    "gamma_type1": dict(mrand=0.2, T=3, gamma_type=1, Acols=600),
}


@pytest.mark.parametrize("name", list(CASES))
def test_bcc_init_run_matches_loop_counts(oracle, name):
    kw = CASES[name]
    text = in_text(**kw)
    S = oracle.State(text)
    T = kw.get("T", 3)
    prob = loopcount.problem_of(S, T)
    S.close()
    nz = 0 if prob[1] is None else prob[1].shape[0]
    n, compared = loopcount.check_run(oracle, text, None, nz, T, kw.get("kappa_squared", 0.0), problem=prob)
    assert n >= 5 and compared >= n // 2


def _libc_draws():
    libc = ctypes.CDLL(ctypes.util.find_library("c"))
    libc.srand(1)
    rmax = 2147483647.0
    return lambda: libc.rand() / rmax


def test_bcc_init_problem_matches_a_second_restatement(oracle):
    """bcc_init's draw order restated in numpy on the C library's rand(): column masks of A
    and T, then A's row masks and bits (column by column over all rows), Z's likewise, T's
    and S's bits with column 0 set, then A_p's cropped draws (maximum / minimum evaluating
    the chosen draw again)."""
    A_rows, M, Z_rows, T, mrnd = 20, 90, 17, 3, 0.25
    text = in_text(mrand=mrnd, A=A_rows, Acols=M, Z=Z_rows, T=T)
    S = oracle.State(text)
    r01 = _libc_draws()

    def masks(n):
        d = np.array([r01() for _ in range(n)])
        return d > mrnd * mrnd, d > mrnd

    def bits(nr, nc):
        X = np.zeros((nr, nc), bool)
        for k in range(nc):
            for r in range(nr):
                X[r, k] = r01() > 0.5
        return X

    Acb, Acj = masks(M)
    masks(T)
    NA, NZ = A_rows - 1, Z_rows - 2
    Arb, Arj = masks(NA)
    A = bits(NA, M)
    Zrb, Zrj = masks(NZ)
    Z = bits(NZ, M)
    Tm = bits(NA, T); Tm[:, 0] = True
    Sm = bits(NZ, T); Sm[:, 0] = True
    pcols = oracle.lib().lko_psize(M)
    Ap = []
    for _ in range(pcols):
        a = r01()
        m1 = 0.9 if 0.9 < a else r01()
        if 0.1 > m1:
            Ap.append(0.1)
        else:
            b = r01()
            Ap.append(0.9 if 0.9 < b else r01())
    assert np.array_equal(S.get("mask_Ac_b").astype(bool), Acb) and np.array_equal(S.get("mask_Ac_j").astype(bool), Acj)
    assert np.array_equal(S.get("mask_Ar_b").astype(bool), Arb) and np.array_equal(S.get("mask_Ar_j").astype(bool), Arj)
    assert np.array_equal(S.get("mask_Zr_b").astype(bool), Zrb) and np.array_equal(S.get("mask_Zr_j").astype(bool), Zrj)
    got = loopcount.problem_of(S, T)
    for g, w in zip(got, (A, Z, Tm, Sm)):
        assert np.array_equal(g, w)
    np.testing.assert_array_equal(S.get("A_p"), np.array(Ap))
    S.close()


@pytest.mark.parametrize("name", ["inactive_rows_cols", "no_controls"])
def test_bcc_init_run_invariants(oracle, name):
    text = in_text(**CASES[name])
    S = oracle.State(text)
    arj, arb = S.get("mask_Ar_j").astype(bool), S.get("mask_Ar_b").astype(bool)
    acj, acb = S.get("mask_Ac_j").astype(bool), S.get("mask_Ac_b").astype(bool)
    S.close()
    r = oracle.run_text(text)
    rows = r.xdrop_a[r.xdrop_a[:, 0] >= 0, 0]
    cols = r.xdrop_a[r.xdrop_a[:, 1] >= 0, 1]
    assert sorted(rows) == sorted(np.flatnonzero(arj)) and sorted(cols) == sorted(np.flatnonzero(acj))
    np.testing.assert_array_equal(r.xdrop_b[r.xdrop_b[:, 0] >= 0, 0], (np.cumsum(arb) - 1)[rows])
    np.testing.assert_array_equal(r.xdrop_b[r.xdrop_b[:, 1] >= 0, 1], (np.cumsum(acb) - 1)[cols])
    assert r.trace[0, 1] == arj.sum() and r.trace[0, 2] == acj.sum()


def test_bcc_init_niter_caps_the_run(oracle):
    full = oracle.run_text(in_text(mrand=0.3))
    three = oracle.run_text(in_text(mrand=0.3, niter=3))
    assert len(three.trace) == 3 and three.trace.tobytes() == full.trace[:3].tobytes()


def test_bcc_sumscores_config_errors_before_any_device_work():
    from lakcluster_c_amd import lkc
    for kw, why in ((dict(Y=32), "GLOBAL_TEST_Y_n_cols"), (dict(typ2="speed"), "GLOBAL_TEST_TYP2"),
                    (dict(nbins=2), "GLOBAL_NBINS")):
        with pytest.raises(lkc.LkcError) as e:
            lkc.bcc_sumscores_test(in_text(**kw), "/tmp/unused_bcc_out")
        assert why in str(e.value), (kw, str(e.value))


# ---- the HIP path ----------------------------------------------------------------------

GPU_CASES = dict(CASES, b24=dict(mrand=0.15, Acols=2500, A=60, Z=50, T=2, B=24),
                 capped=dict(mrand=0.3, niter=3), sparse0=dict(mrand=0.2, TEST_sparse=0))


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(GPU_CASES))
def test_engine_bcc_sumscores_matches_oracle(oracle, tmp_path, name):
    from lakcluster_c_amd import lkc
    text = in_text(**GPU_CASES[name])
    print(f"bcc_sumscores {name}", flush=True)
    a, b = str(tmp_path / "eng"), str(tmp_path / "ora")
    lkc.bcc_sumscores_test(text, a)
    oracle.run_text(text, out_dir=b, write_files=True)
    for f in ("out_trace.txt", "out_xdrop_a.txt", "out_xdrop_b.txt"):
        assert open(f"{a}/{f}").read() == open(f"{b}/{f}").read(), f


@pytest.mark.gpu
def test_engine_bcc_init_handle_and_cli(oracle, tmp_path):
    from lakcluster_c_amd import lkc
    text = in_text(mrand=0.3, Acols=700)
    ref = oracle.run_text(text)
    with lkc.Bcc(text) as run:
        run.run()
        assert run.trace().tobytes() == ref.trace.tobytes()
        xa, xb = run.xdrop()
        np.testing.assert_array_equal(xa, ref.xdrop_a)
        np.testing.assert_array_equal(xb, ref.xdrop_b)
        run.reset()                     # back to the drawn masks, inactive rows included
        run.run()
        assert run.trace().tobytes() == ref.trace.tobytes()
    exe = os.path.join(os.path.dirname(lkc.LIB_PATH), "lkc_driver")
    out = str(tmp_path / "cli")
    r = subprocess.run([exe, out], input=text.encode(), capture_output=True, timeout=300)
    assert r.returncode == 0, r.stderr.decode()
    oracle.run_text(text, out_dir=str(tmp_path / "ora"), write_files=True)
    for f in ("out_trace.txt", "out_xdrop_a.txt", "out_xdrop_b.txt"):
        assert open(f"{out}/{f}").read() == open(str(tmp_path / "ora" / f)).read(), f
